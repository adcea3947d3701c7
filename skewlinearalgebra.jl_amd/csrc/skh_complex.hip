// skh_complex.hip -- kernels of the complex (ComplexF64) skew-Hermitian path: the cooperative
// tridiagonalization (device code in skh_zengine.hpp), the imaginary diagonal of T, and the
// pieces of the explicit Q = H_0 ... H_{n-2} (compact WY with rocBLAS products on the host side).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "skh_fill.hpp"
#include "skh_internal.h"
#include "skh_sync.hpp"
#include "skh_zengine.hpp"
#include "skh_zpanel.hpp"

namespace skh {

// Columns k in [k_begin, k_end) (k_end = n - 1 for the last segment); the norm partials and head
// of column k_begin come from the previous launch, or from zinit when the launch starts with no
// pending reflector (k_begin = 0, or `fresh`: after the blocked panels, with w.Wv zeroed).
__global__ __launch_bounds__(WG, 1) void k_zfused(double2* A, int64_t lda, int64_t n, zc::ZWork w, double2* tau,
                                                  double* eta, unsigned* bar, int* err, int64_t k_begin,
                                                  int64_t k_end, int fresh)
{
    using namespace zc;
    __shared__ ZShared sh;
    __shared__ int s_ok;
    const int G = gridDim.x, g = blockIdx.x;
    unsigned epoch = 0;
    const bool start = k_begin == 0 || fresh;
    if (start) {
        zinit(A, lda, n, w, g, G, k_begin);
        if (!grid_sync_co(bar, ++epoch, err, &s_ok)) return;
    }
    for (int64_t k = k_begin; k < k_end; k++) {
        const ZScal s = zscalars(w, G, sh);
        if (g == 0 && threadIdx.x == 0) {
            st(tau + k, s.tau);
            eta[k] = s.beta;
        }
        if (k == n - 2) {
            // last (length-1) reflector: the factor holds its unit, and the 1x1 trailing block is
            // unchanged by it (|1 - tau| = 1) but still owes the previous reflector's update
            if (g == 0 && threadIdx.x == 0) zlast(A, lda, n, w, s);
            break;
        }
        zpass(A, lda, n, k, k >= 1 && !(start && k == k_begin), s, w, sh, g, G);
        if (!grid_sync_co(bar, ++epoch, err, &s_ok)) return;
        zrows(A, lda, n, k, s, w, sh, g, G, G);
        if (!grid_sync_co(bar, ++epoch, err, &s_ok)) return;
    }
}

constexpr int ZR2K_TR = 32;   // rows of a complex rank-2k wave tile (two row halves per 64 x 64 block)
#ifndef SKH_ZR2K_BJ
#define SKH_ZR2K_BJ 2
#endif
constexpr int ZBJ = SKH_ZR2K_BJ;           // 16-column blocks of a wave tile: 2 (32 x 32) or 4 (32 x 64)
constexpr int ZCPB = 4 / ZBJ;              // wave tiles across a 64 x 64 block
constexpr int ZWPB = 2 * ZCPB;             // wave tiles per 64 x 64 block
static_assert(ZBJ == 2 || ZBJ == 4, "complex rank-2k wave tiles are 32 x 32 or 32 x 64");

// lane offsets of the complex wave tile's operand rows (two 16-row blocks of L, ZBJ of R), apart from
// the real update's tile shape
struct ZR2kRows {
    uint32_t lv[2], lu[2];       // L rows r0 + 16 bi + lane%16, V and W layouts
    uint32_t rv[ZBJ], ru[ZBJ];   // R rows q0 + 16 bj + lane%16
};

// One panel of the blocked complex reduction (device code in skh_zpanel.hpp): co-resident, one
// workgroup per CU, one 64-row block per workgroup (grid >= blocks of the trailing matrix).
__global__ __launch_bounds__(WG, 1) void k_zpanel(double2* A, int64_t lda, int64_t n, int64_t c0, int nbp, zp::ZPWork w,
                                                  double2* tau, double* eta, int* err)
{
    using namespace zp;
    __shared__ ZPShared sh;
    __shared__ cd s_pv[4 * ZKQ * 64], s_pw[4 * ZKQ * 64];
    const int G = gridDim.x, g = blockIdx.x;
    ZRowRegs<ZKQ> rr;
    rr.v = s_pv;
    rr.w = s_pw;
    for (int k = threadIdx.x; k < 4 * ZKQ * 64; k += WG) {
        s_pv[k] = czero();
        s_pw[k] = czero();
    }
    rr.xprev = czero();
    const int64_t nt = cdiv(n, TB), bown = c0 / TB + g;
    {
        const int64_t r = bown * TB + (threadIdx.x & 63);
        rr.xnext = (r < n && r >= c0) ? zc::ld(A + r + c0 * lda) : czero();   // column c0 from its diagonal
    }
    const PanelMap pm = panel_map(n, c0, g, G);
    bool ok = zrow_pass<ZKQ>(A, lda, n, c0, c0, 0, false, true, w, g, G, sh, rr, pm, err);
    for (int i = 0; ok && i < nbp; i++) {
        const int64_t j = c0 + i;
        ok = zphase_b(A, lda, n, c0, j, i, w, sh, tau, eta, g, G, pm, bown < nt && bown >= (j + 1) / TB, err);
        if (ok) ok = zrow_pass<ZKQ>(A, lda, n, c0, j + 1, i + 1, true, i + 1 < nbp, w, g, G, sh, rr, pm, err);
    }
}

// final diagonal entries of the panel columns [c0, c0 + count) into A
__global__ void k_zdfin(double2* A, int64_t lda, const double2* Dfin, int64_t c0, int64_t count)
{
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < count; k += (int64_t)gridDim.x * blockDim.x)
        A[(c0 + k) + (c0 + k) * lda] = Dfin[c0 + k];
}

// Complex skew rank-2k update after a panel (src/hessenberg.jl:158-166, complex case):
// A[r, c] += sum_k L[r, k] conj(R[c, k]) for r >= c >= e, L = [V W], R = [W -V], on the diagonal the
// imaginary part only (the reference adds Omega - Omega^H there).  One wave per 32 x 64 tile, as
// k_rank2k (32 x 32 by default, SKH_ZR2K_BJ), each 16 x 16 block transposed: Re += Rre Lre^T + Rim Lim^T, Im += Rre Lim^T - Rim Lre^T
// (four v_mfma_f64_16x16x4f64 per block and k-step; real and imaginary parts of the interleaved
// operands are the two halves of one 16-byte load).
__device__ __forceinline__ void zr2k_operands(const double2* __restrict__ A, int64_t lda, const double2* __restrict__ U,
                                              int64_t ldu, int64_t c0, int nbp, int nkh, int s, int lq,
                                              const ZR2kRows& o, double2 lf[2], double2 rf[ZBJ])
{
    const int h = s >= nkh;
    const int ks = s - h * nkh;
    const bool ok = ks * 4 + lq < nbp;
    const double2* pv = A + (c0 + ks * 4) * lda;
    const double2* pu = U + (int64_t)ks * 4 * ldu;
    const double2 z = make_double2(0.0, 0.0);
    if (h == 0) {
#pragma unroll
        for (int b = 0; b < 2; b++) lf[b] = ok ? pv[o.lv[b]] : z;
#pragma unroll
        for (int b = 0; b < ZBJ; b++) rf[b] = ok ? pu[o.ru[b]] : z;
    } else {
#pragma unroll
        for (int b = 0; b < 2; b++) lf[b] = ok ? pu[o.lu[b]] : z;
#pragma unroll
        for (int b = 0; b < ZBJ; b++) {
            const double2 t = ok ? pv[o.rv[b]] : z;
            rf[b] = make_double2(-t.x, -t.y);
        }
    }
}

__device__ __forceinline__ void zr2k_mfma(v4d re[2][ZBJ], v4d im[2][ZBJ], const double2 lf[2], const double2 rf[ZBJ])
{
#pragma unroll
    for (int bi = 0; bi < 2; bi++)
#pragma unroll
        for (int bj = 0; bj < ZBJ; bj++) {
            re[bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(rf[bj].x, lf[bi].x, re[bi][bj], 0, 0, 0);
            re[bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(rf[bj].y, lf[bi].y, re[bi][bj], 0, 0, 0);
            im[bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(rf[bj].x, lf[bi].y, im[bi][bj], 0, 0, 0);
            im[bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(-rf[bj].y, lf[bi].x, im[bi][bj], 0, 0, 0);
        }
}

__global__ __launch_bounds__(64, 2) void k_zrank2k(double2* A, int64_t lda, const double2* U, int64_t ldu, int64_t n,
                                                   int64_t e, int64_t c0, int nbp, int64_t b0, int64_t nt,
                                                   int64_t total, int64_t per, FillSegs fs)
{
    const int64_t bb = blockIdx.x;
    if (bb >= 8 * per) {   // workgroups past the update fill the next panel's hand-over slots
        fill_chunk(fs, bb - 8 * per);
        return;
    }
    const int64_t t = (bb % 8) * per + bb / 8;
    if (t >= total) return;
    int64_t ti, tj;
    super_ij(t / ZWPB, nt, ti, tj);
    const int wt = (int)(t % ZWPB);   // row half wt / ZCPB, column part wt % ZCPB of the block
    const int64_t r0 = (b0 + ti) * R2K_BLK + (wt / ZCPB) * ZR2K_TR;
    const int64_t q0 = (b0 + tj) * R2K_BLK + (wt % ZCPB) * (16 * ZBJ);
    if (r0 + ZR2K_TR - 1 < q0) return;   // above the diagonal of a diagonal block: nothing on or below it
    const int lane = threadIdx.x & 63, lc = lane & 15, lq = lane >> 4;
    ZR2kRows o;
#pragma unroll
    for (int b = 0; b < 2; b++) {
        const int64_t rl = min(r0 + b * 16 + lc, n - 1);
        o.lv[b] = (uint32_t)(rl + lq * lda);
        o.lu[b] = (uint32_t)(rl + lq * ldu);
    }
#pragma unroll
    for (int b = 0; b < ZBJ; b++) {
        const int64_t rr = min(q0 + b * 16 + lc, n - 1);
        o.rv[b] = (uint32_t)(rr + lq * lda);
        o.ru[b] = (uint32_t)(rr + lq * ldu);
    }
    double2* C = A + r0 + q0 * lda;
    const uint32_t co = (uint32_t)(lc + lq * lda);
    // strictly below the diagonal and inside the trailing block everywhere: plain loads and stores
    const bool inner = (q0 + 16 * ZBJ - 1 < r0) && (r0 + ZR2K_TR - 1 < n) && (q0 >= e);
    v4d re[2][ZBJ], im[2][ZBJ];
#pragma unroll
    for (int bi = 0; bi < 2; bi++)
#pragma unroll
        for (int bj = 0; bj < ZBJ; bj++) {
            const int64_t row = r0 + lc + bi * 16;
            v4d cr = {0.0, 0.0, 0.0, 0.0}, ci = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int64_t col = q0 + lq + bj * 16 + 4 * q;
                if (inner || (row < n && col >= e && col <= row)) {
                    const double2 x = C[co + bi * 16 + (bj * 16 + 4 * q) * lda];
                    cr[q] = x.x;
                    ci[q] = x.y;
                }
            }
            re[bi][bj] = cr;
            im[bi][bj] = ci;
        }
    const int nkh = (nbp + 3) / 4, S = 2 * nkh;
    double2 la[2], ra[ZBJ], lb[2], rb[ZBJ];
    zr2k_operands(A, lda, U, ldu, c0, nbp, nkh, 0, lq, o, la, ra);
    for (int s = 0; s < S; s += 2) {
        if (s + 1 < S) zr2k_operands(A, lda, U, ldu, c0, nbp, nkh, s + 1, lq, o, lb, rb);
        zr2k_mfma(re, im, la, ra);
        if (s + 1 >= S) break;
        if (s + 2 < S) zr2k_operands(A, lda, U, ldu, c0, nbp, nkh, s + 2, lq, o, la, ra);
        zr2k_mfma(re, im, lb, rb);
    }
#pragma unroll
    for (int bi = 0; bi < 2; bi++) {
        const int64_t row = r0 + lc + bi * 16;
#pragma unroll
        for (int bj = 0; bj < ZBJ; bj++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int64_t col = q0 + lq + bj * 16 + 4 * q;
                double2* pc = C + co + bi * 16 + (bj * 16 + 4 * q) * lda;
                if (inner || (row < n && col >= e && col < row)) *pc = make_double2(re[bi][bj][q], im[bi][bj][q]);
                else if (row < n && col >= e && col == row) reinterpret_cast<double*>(pc)[1] = im[bi][bj][q];
            }
    }
}

// dvim[i] = Im A[i, i]
__global__ void k_zdiag(const double2* A, int64_t lda, int64_t n, double* dvim)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dvim[i] = A[i + i * lda].y;
}

// V (m x nbp, rows c0+1..n-1) of reflectors c0..c0+nbp-1 with explicit unit and zeros above it
__global__ void k_zform_v(const double2* F, int64_t ldf, int64_t n, int64_t c0, int nbp, double2* V, int64_t ldv)
{
    const int k = blockIdx.y;
    const int64_t m = n - c0 - 1;
    for (int64_t rr = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; rr < m; rr += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = c0 + 1 + rr;
        const int64_t head = c0 + k + 1;
        double2 v = make_double2(0.0, 0.0);
        if (r == head) v = make_double2(1.0, 0.0);
        else if (r > head) v = F[r + (c0 + k) * ldf];
        V[rr + k * ldv] = v;
    }
}

// T (nbp x nbp upper) from Gm = V^H V and tau: T[k,k] = tau_k, T[0:k,k] = -tau_k T[0:k,0:k] Gm[0:k,k]
// complex T of the compact WY form (as k_larft): thread i builds row i of T in LDS, no barrier; nbp <= 64
__global__ __launch_bounds__(64) void k_zlarft(const double2* Gm, int64_t ldg, const double2* tau, int nbp, double2* T,
                                               int64_t ldt)
{
    using namespace zc;
    __shared__ double2 row[64 * 64];
    const int i = threadIdx.x;
    if (i >= nbp) return;
    for (int k = 0; k < nbp; k++) {
        cd v = cd{0.0, 0.0};
        const cd tk = ld(tau + k);
        if (i < k) {
            cd acc = cd{0.0, 0.0};
            for (int l = i; l < k; l++) acc = acc + ld(row + l * 64 + i) * ld(Gm + l + k * ldg);
            v = cd{0.0, 0.0} - tk * acc;
        } else if (i == k) {
            v = tk;
        }
        st(row + k * 64 + i, v);
    }
    for (int k = 0; k < nbp; k++) T[i + k * ldt] = row[k * 64 + i];
}

__global__ void k_zeye(double2* Q, int64_t ldq, int64_t n)
{
    const int64_t c = blockIdx.y;
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x)
        Q[r + c * ldq] = make_double2(r == c ? 1.0 : 0.0, 0.0);
}

// Zc[r, j] = i^r C[r, j]: eigenvectors of D^H (iT) D mapped back to iT (D = diag(i^r))
__global__ void k_zphase(const double* C, int64_t n, double2* Zc)
{
    const int64_t j = blockIdx.y;
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
        const double c = C[r + j * n];
        double2 z;
        switch (r & 3) {
            case 0: z = make_double2(c, 0.0); break;
            case 1: z = make_double2(0.0, c); break;
            case 2: z = make_double2(-c, 0.0); break;
            default: z = make_double2(0.0, -c); break;
        }
        Zc[r + j * n] = z;
    }
}

}  // namespace skh

using namespace skh;

extern "C" int skh_internal_zphase(const double* C, int64_t n, double2* Zc, hipStream_t st)
{
    hipLaunchKernelGGL(k_zphase, dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 64)), (unsigned)n),
                       dim3(256), 0, st, C, n, Zc);
    return (int)hipGetLastError();
}

extern "C" int skh_internal_zfused_occupancy(int* blocks_per_cu)
{
    return (int)hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, (const void*)k_zfused, WG, 0);
}

extern "C" int skh_internal_launch_zfused(double2* A, int64_t lda, int64_t n, const zc::ZWork* w, double2* tau,
                                          double* eta, unsigned* bar, int* err, int grid, int64_t k_begin,
                                          int64_t k_end, int fresh, hipStream_t st)
{
    zc::ZWork wc = *w;
    void* args[] = {(void*)&A,   (void*)&lda, (void*)&n,   (void*)&wc,      (void*)&tau,  (void*)&eta,
                    (void*)&bar, (void*)&err, (void*)&k_begin, (void*)&k_end, (void*)&fresh};
    return (int)hipLaunchCooperativeKernel((const void*)k_zfused, dim3(grid), dim3(WG), args, 0, st);
}

extern "C" int skh_internal_zpanel_occupancy(int* blocks_per_cu)
{
    return (int)hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, (const void*)k_zpanel, WG, 0);
}

extern "C" int skh_internal_launch_zpanel(double2* A, int64_t lda, int64_t n, int64_t c0, int nbp, const zp::ZPWork* w,
                                          double2* tau, double* eta, int* err, int grid, hipStream_t st)
{
    zp::ZPWork wc = *w;
    void* args[] = {(void*)&A, (void*)&lda, (void*)&n, (void*)&c0, (void*)&nbp, (void*)&wc, (void*)&tau, (void*)&eta,
                    (void*)&err};
    // barrier-free like the real panel kernel: the same plain launch unless SKH_PANEL_COOPERATIVE=1
    // (the grid is at most one workgroup per CU, the occupancy bound)
    if (skh_internal_panel_cooperative())
        return (int)hipLaunchCooperativeKernel((const void*)k_zpanel, dim3(grid), dim3(WG), args, 0, st);
    return (int)hipLaunchKernel((const void*)k_zpanel, dim3(grid), dim3(WG), args, 0, st);
}

extern "C" int skh_internal_launch_zdfin(double2* A, int64_t lda, const double2* Dfin, int64_t c0, int nbp,
                                         hipStream_t st)
{
    if (nbp <= 0) return 0;
    hipLaunchKernelGGL(k_zdfin, dim3((unsigned)std::min<int64_t>((nbp + 255) / 256, 64)), dim3(256), 0, st, A, lda, Dfin, c0,
                       (int64_t)nbp);
    return (int)hipGetLastError();
}

// trailing update of rows / columns >= e after the panel [c0, c0 + nbp): 64 x 64 tile blocks of the
// lower triangle from block e / 64, two waves each, dealt to the XCDs as k_rank2k; with fs (nseg > 0)
// extra workgroups of the same launch fill those segments with the empty marker
extern "C" int skh_internal_launch_zrank2k(double2* A, int64_t lda, const double2* U, int64_t ldu, int64_t n, int64_t e,
                                           int64_t c0, int nbp, const FillSegs* fs, int nseg, hipStream_t st)
{
    FillSegs f;
    const int64_t nfill = fill_plan(fs, nseg, f);
    if (nfill < 0) return (int)hipErrorInvalidValue;
    const int64_t b0 = e / R2K_BLK;
    const int64_t nt = e < n ? (n + R2K_BLK - 1) / R2K_BLK - b0 : 0;
    const int64_t total = nt * (nt + 1) / 2 * ZWPB;   // wave tiles of the blocks of the triangle
    const int64_t per = (total + 7) / 8;
    if (per * 8 + nfill == 0) return 0;
    hipLaunchKernelGGL(k_zrank2k, dim3((unsigned)(per * 8 + nfill)), dim3(64), 0, st, A, lda, U, ldu, n, e, c0, nbp, b0, nt,
                       total, per, f);
    return (int)hipGetLastError();
}

extern "C" int skh_internal_zdiag(const double2* A, int64_t lda, int64_t n, double* dvim, hipStream_t st)
{
    hipLaunchKernelGGL(k_zdiag, dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 1024))), dim3(256),
                       0, st, A, lda, n, dvim);
    return (int)hipGetLastError();
}

extern "C" int skh_internal_zform_v(const double2* F, int64_t ldf, int64_t n, int64_t c0, int nbp, double2* V,
                                    int64_t ldv, hipStream_t st)
{
    const int64_t m = n - c0 - 1;
    hipLaunchKernelGGL(k_zform_v, dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>((m + 255) / 256, 64)), nbp),
                       dim3(256), 0, st, F, ldf, n, c0, nbp, V, ldv);
    return (int)hipGetLastError();
}

extern "C" int skh_internal_zlarft(const double2* Gm, int64_t ldg, const double2* tau, int nbp, double2* T, int64_t ldt,
                                   hipStream_t st)
{
    if (nbp > 64) return (int)hipErrorInvalidValue;
    hipLaunchKernelGGL(k_zlarft, dim3(1), dim3(64), 0, st, Gm, ldg, tau, nbp, T, ldt);
    return (int)hipGetLastError();
}

extern "C" int skh_internal_zeye(double2* Q, int64_t ldq, int64_t n, hipStream_t st)
{
    hipLaunchKernelGGL(k_zeye, dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 64)), (unsigned)n),
                       dim3(256), 0, st, Q, ldq, n);
    return (int)hipGetLastError();
}
