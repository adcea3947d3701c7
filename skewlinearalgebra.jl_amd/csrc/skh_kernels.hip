// skh_kernels.hip -- gfx950 kernels of the skew tridiagonalization and the engines that drive them.
//
//  * k_panel     : one panel of the blocked reduction, a persistent kernel over the whole chip
//                  whose workgroups hand values over through polled slots (no grid barrier),
//                  replacing skewlatrd! (src/hessenberg.jl:92-131) for one matrix.
//  * k_rank2k    : skew rank-2k trailing update on FP64 MFMA (src/hessenberg.jl:158-166),
//                  lower triangle only, no Omega buffer.
//  * Q formation and the batched gradient helpers (compact WY with rocBLAS products).
// The compact engine for batches of small matrices is in skh_compact.hip.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <algorithm>

#include "skh_device.hpp"
#include "skh_fill.hpp"
#include "skh_internal.h"
#include "skh_sync.hpp"

namespace skh {

// Optional phase clock (workgroup 0, 100 MHz wall clock), accumulated over all launches:
// ph[2] phase B (poll of the row pass's outputs, scalars, t, matvec), ph[3] row passes (with the
// poll of the matvec partials); inside them ph[4..7] the row pass (loads and poll, correction of
// w, column update with its partials, hand-over to the next column) and phase B: ph[11] tile
// range and first tile, ph[12] poll, ph[8] scalars, ph[9] t, ph[10] matvec.  ph[0] and ph[1] are
// the time between the phases (no grid barrier any more: both stay near zero).
// panels up to 64 wide fit two workgroups per CU (<= 256 registers per lane); 128 wide takes one.
// MULTI: the grid has fewer workgroups than the panel has 64-row blocks, so a workgroup also
// processes blocks other than its own (rows of V and W then come from memory); the common
// instance (one block per workgroup) carries none of that path.
// The launch has no grid barrier: each workgroup polls the hand-over slots of its inputs (see
// Work).  All workgroups are co-resident (the grid is at most the occupancy bound; see
// skh_internal_launch_panel), and every poll is bounded: a lost producer flags `err` and ends the
// launch.
template <int KQ, bool MULTI>
__global__ __launch_bounds__(WG, (KQ <= 16) ? 2 : 1) void k_panel(double* A, int64_t lda, int64_t n, int64_t c0, int nbp, Work w,
                                              double* tau, double* eta, int* err, unsigned long long* ph)
{
    // the 32-wide instance keeps each rank's first matvec tile in LDS (32 KB, room made by the
    // 4-row flush buffer: two workgroups per CU still fit the 160 KB); SKH_PANEL_RESIDENT=0 builds the
    // form without it (16-row flush buffer)
#ifndef SKH_PANEL_RESIDENT
#define SKH_PANEL_RESIDENT 1
#endif
    constexpr bool RES = SKH_PANEL_RESIDENT && KQ == 8;
    constexpr int MROWS = (KQ == 8 && !RES) ? 16 : 4;
    __shared__ RowShared<MROWS> sh;
    __shared__ double2 s_rt[RES ? 8 * WG : 1];
    __shared__ unsigned long long s_clk[PH_SLOTS];
    const int G = gridDim.x, g = blockIdx.x;
    unsigned long long* clk = ((ph != nullptr) && g == 0 && threadIdx.x == 0) ? s_clk : nullptr;
    if (clk) {
        for (int k = 0; k < PH_SLOTS; k++) s_clk[k] = 0;
        clk_begin(clk);
        clk[PH_OUTER] = clk[PH_INNER];
    }
    __shared__ double s_pv[4 * KQ * 64], s_pw[4 * KQ * 64];
    RowRegs<KQ> rr;
    rr.v = s_pv;
    rr.w = s_pw;
    for (int k = threadIdx.x; k < 4 * KQ * 64; k += WG) {
        s_pv[k] = 0.0;
        s_pw[k] = 0.0;
    }
    rr.xprev = 0.0;
    {
        const int64_t r = (c0 / TB + g) * TB + (threadIdx.x & 63);
        rr.xnext = (r < n && r > c0) ? A[r + c0 * lda] : 0.0;
    }
    const PanelMap pm = panel_map(n, c0, g, G);
    if constexpr (RES) tile_keep(A, lda, n, pm, g, s_rt);
    // last 64-row block this workgroup finalises rows of (-1: none)
    const int64_t nt = cdiv(n, TB), bown = c0 / TB + g;
    const int64_t blast = (bown < nt) ? (MULTI ? bown + (int64_t)G * ((nt - 1 - bown) / G) : bown) : -1;
    bool ok = row_pass<KQ, MULTI, MROWS>(A, lda, n, c0, c0, 0, false, true, w, g, G, sh, rr, pm, err, clk);
    for (int i = 0; ok && i < nbp; i++) {
        const int64_t j = c0 + i;
        clk_outer(clk, 0);
        ok = phase_b<MROWS, RES>(A, lda, n, c0, j, i, w, sh, tau, eta, g, G, pm, blast >= (j + 1) / TB, s_rt, err, clk);
        clk_outer(clk, 2);
        if (ok) ok = row_pass<KQ, MULTI, MROWS>(A, lda, n, c0, j + 1, i + 1, true, i + 1 < nbp, w, g, G, sh, rr, pm, err, clk);
        clk_outer(clk, 3);
    }
    if (clk) {
        clk_outer(clk, 0);
        for (int k = 0; k < 13; k++) ph[k] += s_clk[k];
    }
}

// One wave per R2K_TR x R2K_TC tile; tile blocks (64x64, R2K_WPB tiles each) of the trailing triangle
// from block b0 = e/64, consecutive indices on one XCD (blocks are dealt to XCDs round robin).
// Workgroups from 8 per on fill fs (the next panel's hand-over slots) instead.
__global__ __launch_bounds__(64, 2) void k_rank2k(double* A, int64_t lda, const double* U, int64_t ldu, int64_t n,
                                                  int64_t e, int64_t c0, int nbp, int64_t b0, int64_t nt,
                                                  int64_t total, int64_t per, FillSegs fs)
{
    const int64_t b = blockIdx.x;
    if (b >= 8 * per) {
        fill_chunk(fs, b - 8 * per);
        return;
    }
    const int64_t t = (b % 8) * per + b / 8;
    if (t >= total) return;
    int64_t i, j;
    super_ij(t / R2K_WPB, nt, i, j);
    const int wt = (int)(t % R2K_WPB);   // wave tile inside the block: row-major over (TR x TC) sub-tiles
    const int64_t r0 = (b0 + i) * R2K_BLK + (wt / (R2K_BLK / R2K_TC)) * R2K_TR;
    const int64_t q0 = (b0 + j) * R2K_BLK + (wt % (R2K_BLK / R2K_TC)) * R2K_TC;
    if (r0 + R2K_TR - 1 <= q0) return;   // the upper half of a diagonal block: nothing below the diagonal
    rank2k_wave(A, lda, U, ldu, n, e, c0, nbp, r0, q0);
}

__global__ void k_finish(double* A, int64_t lda, int64_t n, double* tau, double* eta)
{
    if (threadIdx.x == 0 && n >= 2) {
        eta[n - 2] = A[(n - 1) + (n - 2) * lda];
        tau[n - 2] = 0.0;
    }
}

// fill entries [lo, hi) of `rows` rows (ld doubles apart) at each segment's base with the empty marker
__global__ void k_fill_empty(FillSegs fs)
{
    const double e = empty_slot();
    const int sgi = blockIdx.z;
    const int64_t row = blockIdx.y;
    if (row >= fs.rows[sgi]) return;
    double* s = fs.p[sgi] + row * fs.ld[sgi];
    for (int64_t r = fs.lo[sgi] + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < fs.hi[sgi];
         r += (int64_t)gridDim.x * blockDim.x)
        s[r] = e;
}

__global__ void k_zero_u32(unsigned* p, int count)
{
    for (int k = threadIdx.x; k < count; k += blockDim.x) p[k] = 0u;
}

// ---------------------------------------------------------------------------------------
// Explicit Q = H_0 ... H_{n-2} from the 'L'-layout factors (LAPACK orgtr('L') convention:
// reflector k has v[k+1] = 1 implicit, v[k+2:] = F[k+2:, k]).  Blocked compact WY from the
// last block to the first: Q22 <- Q22 - V (T (V^T Q22)), T from larft('F','C').
// ---------------------------------------------------------------------------------------
__global__ void k_eye(double* Q, int64_t ldq, int64_t n)
{
    const int64_t c = blockIdx.y;
    for (int64_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x)
        Q[r + c * ldq] = (r == c) ? 1.0 : 0.0;
}

// V (m x nbp, rows c0+1..n-1) with the implicit unit and explicit zeros above it
__global__ void k_form_v(const double* F, int64_t ldf, int64_t n, int64_t c0, int nbp, double* V, int64_t ldv)
{
    const int k = blockIdx.y;
    const int64_t m = n - c0 - 1;
    for (int64_t rr = blockIdx.x * blockDim.x + threadIdx.x; rr < m; rr += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = c0 + 1 + rr;
        const int64_t head = c0 + k + 1;
        double v = 0.0;
        if (r == head) v = 1.0;
        else if (r > head) v = F[r + (c0 + k) * ldf];
        V[rr + k * ldv] = v;
    }
}

// T (nbp x nbp, upper) from the Gram matrix Gm = V^T V and tau:
// T[k,k] = tau_k, T[0:k,k] = -tau_k T[0:k,0:k] Gm[0:k,k]   (one workgroup)
// T (nbp x nbp, upper triangular, ldt) of the compact WY form I - V T V^T from Gm = V^T V and tau (forward,
// columnwise): T[k, k] = tau[k], T[i, k] = -tau[k] sum_{l=i}^{k-1} T[i, l] Gm[l, k] for i < k, 0 below.
// Entry (i, k) needs only row i of T and column k of Gm, so thread i builds row i alone: the row stays in
// LDS (column-major by thread, no bank conflicts), no barrier is needed, and the Gm loads of one step are
// independent of each other (coalesced across threads).  nbp <= LARFT_MAX.
constexpr int LARFT_MAX = 64;

__device__ __forceinline__ void larft_rows(const double* __restrict__ g, int64_t ldg, const double* __restrict__ tau,
                                           int nbp, double* __restrict__ t, int64_t ldt, double* row)
{
    const int i = threadIdx.x;
    if (i >= nbp) return;
    for (int k = 0; k < nbp; k++) {
        double v = 0.0;
        if (i < k) {
            double acc = 0.0;
            for (int l = i; l < k; l++) acc += row[l * LARFT_MAX + i] * g[l + k * ldg];
            v = -tau[k] * acc;
        } else if (i == k) {
            v = tau[k];
        }
        row[k * LARFT_MAX + i] = v;
    }
    for (int k = 0; k < nbp; k++) t[i + k * ldt] = row[k * LARFT_MAX + i];
}

__global__ __launch_bounds__(LARFT_MAX) void k_larft(const double* Gm, int64_t ldg, const double* tau, int nbp, double* T,
                                                     int64_t ldt)
{
    __shared__ double row[LARFT_MAX * LARFT_MAX];
    larft_rows(Gm, ldg, tau, nbp, T, ldt, row);
}

// ---------------------------------------------------------------------------------------
// Batched pieces of the log|Pf| gradient  d log|Pf(A)| / dA = -1/2 A^{-1},
//   A^{-1} = Q T^{-1} Q^T,  T^{-1} = P^T [[0, -C^{-T}], [C^{-1}, 0]] P,
// with C (n/2 x n/2, lower bidiagonal) C[a,a] = -eta[2a], C[a,a-1] = eta[2a-1] the odd/even
// coupling block of T.  Writing Q_O = Q[:, 0::2], Q_E = Q[:, 1::2] and K = Q_E C^{-1} Q_O^T,
// A^{-1} = K - K^T.  Q comes from the factors (compact WY, library GEMMs); C^{-1} Q_O^T is a
// forward substitution, K one GEMM.
// ---------------------------------------------------------------------------------------
__global__ void k_eye_batched(double* const* Q, int64_t n)
{
    double* q = Q[blockIdx.z];
    const int64_t c = blockIdx.y;
    for (int64_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x)
        q[r + c * n] = (r == c) ? 1.0 : 0.0;
}

__global__ void k_form_v_batched(const double* const* F, int64_t n, int64_t c0, int nbp, double* V, int64_t vstride)
{
    const int k = blockIdx.y;
    const double* f = F[blockIdx.z];
    double* v = V + (int64_t)blockIdx.z * vstride;
    const int64_t m = n - c0 - 1;
    for (int64_t rr = blockIdx.x * blockDim.x + threadIdx.x; rr < m; rr += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = c0 + 1 + rr, head = c0 + k + 1;
        v[rr + k * m] = (r == head) ? 1.0 : ((r > head) ? f[r + (c0 + k) * n] : 0.0);
    }
}

__global__ __launch_bounds__(LARFT_MAX) void k_larft_batched(const double* Gm, int64_t gstride, const double* taueta,
                                                             const int64_t* toffs, int64_t c0, int nbp, double* T,
                                                             int64_t tstride)
{
    __shared__ double row[LARFT_MAX * LARFT_MAX];
    const int z = blockIdx.x;
    larft_rows(Gm + (int64_t)z * gstride, nbp, taueta + toffs[z] + c0, nbp, T + (int64_t)z * tstride, nbp, row);
}

// Zt (n x n/2) = (C^{-1} Q_O^T)^T: for each row c of Q, forward substitution along a.
__global__ void k_bidiag_solve(double* const* Q, const double* taueta, const int64_t* toffs, int64_t n, double* Zt,
                               int64_t zstride)
{
    const int z = blockIdx.y;
    const double* q = Q[z];
    const double* eta = taueta + toffs[z] + n;
    double* zt = Zt + (int64_t)z * zstride;
    const int64_t h = n / 2;
    for (int64_t c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
        double prev = 0.0;
        for (int64_t a = 0; a < h; a++) {
            const double rhs = q[c + (2 * a) * n] - ((a > 0) ? eta[2 * a - 1] * prev : 0.0);
            prev = rhs / (-eta[2 * a]);
            zt[c + a * n] = prev;
        }
    }
}

// G = alpha (K - K^T) through 32x32 LDS tiles (grid.z = matrix)
__global__ void k_skew_out(const double* K, int64_t kstride, double* const* G, int64_t n, double alpha)
{
    __shared__ double tile[32][33];
    const double* k = K + (int64_t)blockIdx.z * kstride;
    double* g = G[blockIdx.z];
    const int64_t r0 = (int64_t)blockIdx.x * 32, c0 = (int64_t)blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8 threads
    for (int yy = ty; yy < 32; yy += 8) {
        const int64_t r = c0 + tx, c = r0 + yy;   // K^T block: K[c0 + tx, r0 + yy]
        tile[yy][tx] = (r < n && c < n) ? k[r + c * n] : 0.0;
    }
    __syncthreads();
    for (int yy = ty; yy < 32; yy += 8) {
        const int64_t r = r0 + tx, c = c0 + yy;
        if (r < n && c < n) g[r + c * n] = alpha * (k[r + c * n] - tile[tx][yy]);
    }
}

// v_mfma_f64_16x16x4f64 lane-layout probe (diagnostics): D = A*B for three operand pairs.
__global__ void k_mfma_probe(double* out)
{
    const int l = threadIdx.x;
    const int i = l & 15, k = l >> 4;
    v4d z = {0.0, 0.0, 0.0, 0.0};
    // 1: A[i][k] = i + 16k, B[k][j] = (k == 0)  -> D[i][j] = i
    v4d d1 = __builtin_amdgcn_mfma_f64_16x16x4f64((double)(i + 16 * k), (k == 0) ? 1.0 : 0.0, z, 0, 0, 0);
    // 2: A[i][k] = (k == 0), B[k][j] = j + 16k -> D[i][j] = j
    v4d d2 = __builtin_amdgcn_mfma_f64_16x16x4f64((k == 0) ? 1.0 : 0.0, (double)(i + 16 * k), z, 0, 0, 0);
    // 3: A[i][k] = sin(i*4+k), B[k][j] = cos(k*16+j): general product
    v4d d3 = __builtin_amdgcn_mfma_f64_16x16x4f64(sin((double)(i * 4 + k)), cos((double)(k * 16 + i)), z, 0, 0, 0);
    for (int q = 0; q < 4; q++) {
        out[l * 4 + q] = d1[q];
        out[256 + l * 4 + q] = d2[q];
        out[512 + l * 4 + q] = d3[q];
    }
}

}  // namespace skh

// =========================================================================================
// host side: launches (device pointers, caller's stream)
// =========================================================================================
using namespace skh;

static const void* panel_kernel_for(int nbmax, int multi)
{
    if (multi) {
        if (nbmax <= 32) return (const void*)k_panel<8, true>;
        if (nbmax <= 64) return (const void*)k_panel<16, true>;
        return (const void*)k_panel<32, true>;
    }
    if (nbmax <= 32) return (const void*)k_panel<8, false>;
    if (nbmax <= 64) return (const void*)k_panel<16, false>;
    return (const void*)k_panel<32, false>;
}

// The panel's workgroups wait on each other through bounded polls, so all of them must be resident
// at once.  The grid never exceeds the occupancy bound (blocks per CU x CUs) and launches on one
// stream run in order, so on a device that runs nothing beside the factorization a plain launch
// places every workgroup at once, and it costs about 22 us less per panel than the cooperative one
// (n = 8192, 256 panels: 212.4 -> 206.7 ms on MI355X).  SKH_PANEL_COOPERATIVE=1 takes the cooperative
// launch, whose co-residency the runtime guarantees also beside other work on the device.  A plain
// launch that does not become resident in time ends with the bounded polls' error, not a hang.
extern "C" int skh_internal_panel_cooperative()
{
    const char* e = getenv("SKH_PANEL_COOPERATIVE");
    return (e && atoi(e) != 0) ? 1 : 0;
}

extern "C" int skh_internal_launch_panel(const PanelLaunch* pl, hipStream_t st)
{
    void* args[] = {(void*)&pl->A,   (void*)&pl->lda, (void*)&pl->n,   (void*)&pl->c0,  (void*)&pl->nbp,
                    (void*)&pl->w,   (void*)&pl->tau, (void*)&pl->eta, (void*)&pl->err, (void*)&pl->ph};
    const void* k = panel_kernel_for(pl->nbmax, pl->multi);
    if (skh_internal_panel_cooperative())
        return (int)hipLaunchCooperativeKernel(k, dim3(pl->grid), dim3(WG), args, 0, st);
    return (int)hipLaunchKernel(k, dim3(pl->grid), dim3(WG), args, 0, st);
}

extern "C" int skh_internal_fill_empty(const FillSegs* fs, int nseg, hipStream_t st)
{
    if (nseg < 0 || nseg > FillSegs::MAXSEG) return (int)hipErrorInvalidValue;
    if (nseg == 0) return 0;
    FillSegs f = *fs;
    int64_t ymax = 1, cnt = 1;
    for (int k = 0; k < nseg; k++) {
        ymax = std::max(ymax, f.rows[k]);
        cnt = std::max(cnt, f.hi[k] - f.lo[k]);
    }
    for (int k = nseg; k < FillSegs::MAXSEG; k++) f.rows[k] = 0;
    const unsigned gx = (unsigned)std::max<int64_t>(1, std::min<int64_t>((cnt + 255) / 256, 16));
    hipLaunchKernelGGL(k_fill_empty, dim3(gx, (unsigned)ymax, (unsigned)nseg), dim3(256), 0, st, f);
    return (int)hipGetLastError();
}

extern "C" int skh_internal_panel_occupancy(int nbmax, int* blocks_per_cu)
{
    int a = 0, b = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, panel_kernel_for(nbmax, 0), WG, 0);
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, panel_kernel_for(nbmax, 1), WG, 0);
    *blocks_per_cu = a < b ? a : b;
    return (int)e;
}

extern "C" int skh_internal_launch_rank2k(double* A, int64_t lda, const double* U, int64_t ldu, int64_t n, int64_t e,
                                          int64_t c0, int nbp, const FillSegs* fs, int nseg, hipStream_t st)
{
    FillSegs f;
    const int64_t nfill = fill_plan(fs, nseg, f);
    if (nfill < 0) return (int)hipErrorInvalidValue;
    const int64_t b0 = e / R2K_BLK, nt = e < n ? (n + R2K_BLK - 1) / R2K_BLK - b0 : 0;
    const int64_t total = nt * (nt + 1) / 2 * R2K_WPB;   // wave tiles of the blocks of the triangle
    const int64_t per = (total + 7) / 8;
    if (per * 8 + nfill == 0) return 0;
    hipLaunchKernelGGL(k_rank2k, dim3((unsigned)(per * 8 + nfill)), dim3(64), 0, st, A, lda, U, ldu, n, e, c0, nbp, b0,
                       nt, total, per, f);
    return (int)hipGetLastError();
}

extern "C" int skh_internal_launch_finish(double* A, int64_t lda, int64_t n, double* tau, double* eta, hipStream_t st)
{
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(64), 0, st, A, lda, n, tau, eta);
    return (int)hipGetLastError();
}

extern "C" int skh_internal_zero_u32(unsigned* p, int count, hipStream_t st)
{
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(256), 0, st, p, count);
    return (int)hipGetLastError();
}

extern "C" int skh_internal_mfma_probe(double* dout, hipStream_t st)
{
    hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, st, dout);
    return (int)hipGetLastError();
}

extern "C" int skh_internal_eye(double* Q, int64_t ldq, int64_t n, hipStream_t st)
{
    hipLaunchKernelGGL(k_eye, dim3((unsigned)std::min<int64_t>((n + 255) / 256, 64), (unsigned)n), dim3(256), 0, st, Q,
                       ldq, n);
    return (int)hipGetLastError();
}

extern "C" int skh_internal_form_v(const double* F, int64_t ldf, int64_t n, int64_t c0, int nbp, double* V, int64_t ldv,
                                   hipStream_t st)
{
    const int64_t m = n - c0 - 1;
    hipLaunchKernelGGL(k_form_v, dim3((unsigned)std::min<int64_t>((m + 255) / 256, 64), (unsigned)nbp), dim3(256), 0,
                       st, F, ldf, n, c0, nbp, V, ldv);
    return (int)hipGetLastError();
}

extern "C" int skh_internal_larft(const double* Gm, int64_t ldg, const double* tau, int nbp, double* T, int64_t ldt,
                                  hipStream_t st)
{
    if (nbp > LARFT_MAX) return (int)hipErrorInvalidValue;
    hipLaunchKernelGGL(k_larft, dim3(1), dim3(LARFT_MAX), 0, st, Gm, ldg, tau, nbp, T, ldt);
    return (int)hipGetLastError();
}

extern "C" int skh_internal_eye_batched(double* const* Q, int64_t n, int count, hipStream_t st)
{
    hipLaunchKernelGGL(k_eye_batched, dim3((unsigned)std::min<int64_t>((n + 255) / 256, 16), (unsigned)n, (unsigned)count),
                       dim3(256), 0, st, Q, n);
    return (int)hipGetLastError();
}

extern "C" int skh_internal_form_v_batched(const double* const* F, const int64_t* /*unused*/, int64_t n, int64_t c0,
                                           int nbp, double* V, int64_t vstride, int count, hipStream_t st)
{
    const int64_t m = n - c0 - 1;
    hipLaunchKernelGGL(k_form_v_batched,
                       dim3((unsigned)std::min<int64_t>((m + 255) / 256, 16), (unsigned)nbp, (unsigned)count), dim3(256),
                       0, st, F, n, c0, nbp, V, vstride);
    return (int)hipGetLastError();
}

extern "C" int skh_internal_larft_batched(const double* Gm, int64_t gstride, const double* taueta, const int64_t* toffs,
                                          int64_t c0, int nbp, double* T, int64_t tstride, int count, hipStream_t st)
{
    if (nbp > LARFT_MAX) return (int)hipErrorInvalidValue;
    hipLaunchKernelGGL(k_larft_batched, dim3((unsigned)count), dim3(LARFT_MAX), 0, st, Gm, gstride, taueta, toffs, c0,
                       nbp, T, tstride);
    return (int)hipGetLastError();
}

extern "C" int skh_internal_bidiag_solve(double* const* Q, const double* taueta, const int64_t* toffs, int64_t n,
                                         double* Zt, int64_t zstride, int count, hipStream_t st)
{
    hipLaunchKernelGGL(k_bidiag_solve, dim3((unsigned)std::max<int64_t>(1, (n + 63) / 64), (unsigned)count), dim3(64), 0,
                       st, Q, taueta, toffs, n, Zt, zstride);
    return (int)hipGetLastError();
}

extern "C" int skh_internal_skew_out(const double* K, int64_t kstride, double* const* G, int64_t n, double alpha,
                                     int count, hipStream_t st)
{
    const unsigned t = (unsigned)((n + 31) / 32);
    hipLaunchKernelGGL(k_skew_out, dim3(t, t, (unsigned)count), dim3(256), 0, st, K, kstride, G, n, alpha);
    return (int)hipGetLastError();
}
