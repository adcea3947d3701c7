// skh_kernels.hip -- gfx950 kernels of the skew tridiagonalization and the engines that drive them.
//
//  * k_panel     : one panel of the blocked reduction, a cooperative persistent kernel over the
//                  whole chip (two grid barriers per column), replacing skewlatrd!
//                  (src/hessenberg.jl:92-131) for one matrix.
//  * k_rank2k    : skew rank-2k trailing update on FP64 MFMA (src/hessenberg.jl:158-166),
//                  lower triangle only, no Omega buffer.
//  * k_batched   : one workgroup per matrix (work queue over a size-sorted batch): the same
//                  panel phases with G = 1, the MFMA update, then log|Pf| and sign
//                  (src/pfaffian.jl:167-179) and optionally the gradient -1/2 A^{-1}.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <algorithm>

#include "skh_device.hpp"
#include "skh_internal.h"

namespace skh {

// ---------------------------------------------------------------------------------------
// grid barrier for the cooperative panel kernel.  The counter only grows within one
// factorization; the host passes the value it had when the launch started.  A bounded spin
// turns a lost workgroup into an error flag instead of a hang.
// ---------------------------------------------------------------------------------------
constexpr unsigned SPIN_LIMIT = 1u << 26;

__device__ bool grid_sync(unsigned* bar, unsigned target, int* err, int* s_ok)
{
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        atomicAdd(bar, 1u);
        unsigned spins = 0;
        int ok = 1;
        while (__hip_atomic_load(bar, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
            __builtin_amdgcn_s_sleep(1);
            if (++spins > SPIN_LIMIT) {
                atomicOr(err, 1);
                ok = 0;
                break;
            }
        }
        if (ok && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) ok = 0;
        __threadfence();
        *s_ok = ok;
    }
    __syncthreads();
    return *s_ok != 0;
}

__global__ __launch_bounds__(WG) void k_panel(double* A, int64_t lda, int64_t n, int64_t c0, int nbp, Work w,
                                              double* tau, double* eta, unsigned* bar, unsigned base, int* err)
{
    __shared__ RowShared sh;
    __shared__ int s_ok;
    const int G = gridDim.x, g = blockIdx.x;
    unsigned target = base;
    row_pass(A, lda, n, c0, c0, 0, false, true, w, g, G, sh);
    for (int i = 0; i < nbp; i++) {
        const int64_t j = c0 + i;
        target += G;
        if (!grid_sync(bar, target, err, &s_ok)) return;
        column_scalars(A, lda, n, j, w, sh, tau, eta, g == 0);
        reduce_t(A, lda, n, c0, j, i, w, sh, g, G);
        skew_matvec(A, lda, n, j, w, sh, g, G);
        target += G;
        if (!grid_sync(bar, target, err, &s_ok)) return;
        row_pass(A, lda, n, c0, j + 1, i + 1, true, i + 1 < nbp, w, g, G, sh);
    }
}

__global__ __launch_bounds__(WG) void k_rank2k(double* A, int64_t lda, const double* U, int64_t ldu, int64_t n,
                                               int64_t rlo, int64_t clo, int64_t c0, int nbp, TileGrid tg,
                                               int64_t per)
{
    __shared__ MfmaShared sm;
    const int64_t b = blockIdx.x;
    const int64_t t = (b % 8) * per + b / 8;   // consecutive tiles share an XCD (and its L2)
    if (t >= tg.total) return;
    int64_t r0, q0;
    tile_coords(tg, t, r0, q0);
    rank2k_tile(A, lda, U, ldu, n, rlo, clo, c0, nbp, r0, q0, sm);
}

__global__ void k_finish(double* A, int64_t lda, int64_t n, double* tau, double* eta)
{
    if (threadIdx.x == 0 && n >= 2) {
        eta[n - 2] = A[(n - 1) + (n - 2) * lda];
        tau[n - 2] = 0.0;
    }
}

__global__ void k_zero_u32(unsigned* p, int count)
{
    for (int k = threadIdx.x; k < count; k += blockDim.x) p[k] = 0u;
}

// ---------------------------------------------------------------------------------------
// One-workgroup tridiagonalization of an n x n matrix M (ld even), for the batched engine.
// ---------------------------------------------------------------------------------------
__device__ void tridiag_team1(double* M, int64_t ld, int64_t n, int nb, const Work& w, double* tau, double* eta,
                              RowShared& sh, MfmaShared& sm)
{
    for (int64_t c0 = 0; c0 <= n - 3; c0 += nb) {
        const int nbp = (int)min((int64_t)nb, n - 2 - c0);
        row_pass(M, ld, n, c0, c0, 0, false, true, w, 0, 1, sh);
        for (int i = 0; i < nbp; i++) {
            const int64_t j = c0 + i;
            __syncthreads();
            column_scalars(M, ld, n, j, w, sh, tau, eta, true);
            reduce_t(M, ld, n, c0, j, i, w, sh, 0, 1);
            skew_matvec(M, ld, n, j, w, sh, 0, 1);
            __syncthreads();
            row_pass(M, ld, n, c0, j + 1, i + 1, true, i + 1 < nbp, w, 0, 1, sh);
        }
        __syncthreads();
        const int64_t e = c0 + nbp;
        const TileGrid tg = tile_grid(n, e, e);
        for (int64_t t = 0; t < tg.total; t++) {
            int64_t r0, q0;
            tile_coords(tg, t, r0, q0);
            rank2k_tile(M, ld, w.W, w.ldw, n, e, e, c0, nbp, r0, q0, sm);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0 && n >= 2) {
        eta[n - 2] = M[(n - 1) + (n - 2) * ld];
        tau[n - 2] = 0.0;
    }
    __syncthreads();
}

// log|Pf| and sign from eta, tau of an even-n factorization (src/pfaffian.jl:170-178):
// logpf = sum_{odd i} log|ev_i|, sign = prod sign(-ev_i) * (-1)^{#tau != 0}.
__device__ void logpf_reduce(const double* eta, const double* tau, int64_t n, double* out_lp, double* out_sg)
{
    __shared__ double s_l[4];
    __shared__ int s_c[4][3];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    double l = 0.0;
    int nneg = 0, nzero = 0, ntau = 0;
    for (int64_t i = tid; i < n - 1; i += WG) {
        if ((i & 1) == 0) {
            const double e = eta[i];
            l += log(fabs(e));
            nneg += (e > 0.0);   // sign(-e) < 0
            nzero += (e == 0.0);
        }
        ntau += (tau[i] != 0.0);
    }
    l = wave_sum(l);
    double dn = wave_sum((double)nneg), dz = wave_sum((double)nzero), dt = wave_sum((double)ntau);
    if (lane == 0) {
        s_l[wave] = l;
        s_c[wave][0] = (int)dn;
        s_c[wave][1] = (int)dz;
        s_c[wave][2] = (int)dt;
    }
    __syncthreads();
    if (tid == 0) {
        double L = s_l[0] + s_l[1] + s_l[2] + s_l[3];
        int cn = s_c[0][0] + s_c[1][0] + s_c[2][0] + s_c[3][0];
        int cz = s_c[0][1] + s_c[1][1] + s_c[2][1] + s_c[3][1];
        int ct = s_c[0][2] + s_c[1][2] + s_c[2][2] + s_c[3][2];
        *out_lp = L;
        *out_sg = (cz > 0) ? 0.0 : (((cn + ct) & 1) ? -1.0 : 1.0);
    }
    __syncthreads();
}

__global__ __launch_bounds__(WG) void k_batched(BatchArgs a)
{
    __shared__ RowShared sh;
    __shared__ MfmaShared sm;
    __shared__ int64_t s_idx;
    const int tid = threadIdx.x;
    const int g = blockIdx.x;
    double* slot = a.slots + (int64_t)g * a.slot_stride;
    for (;;) {
        if (tid == 0) s_idx = (int64_t)atomicAdd(a.queue, 1u);
        __syncthreads();
        const int64_t idx = s_idx;
        __syncthreads();
        if (idx >= a.batch) break;
        const int64_t b = a.order[idx];
        const int64_t n = a.ns[b];
        if (n % 2 == 1) {
            if (tid == 0) {
                a.logpf[b] = -INFINITY;
                a.sgn[b] = 0.0;
            }
            continue;
        }
        if (n == 0) {
            if (tid == 0) {
                a.logpf[b] = 0.0;
                a.sgn[b] = 1.0;
            }
            continue;
        }
        const int64_t ld = (n + 1) & ~int64_t(1);
        // slot layout: M (ld x n) | W (ld x nb) | tau, eta (n) | Pnorm | PS | Tv | Prow | Pcol
        double* M = slot;
        Work w;
        w.W = M + ld * n;
        w.ldw = ld;
        double* tau = w.W + ld * a.nb;
        double* eta = tau + ld;
        const int64_t nt = cdiv(n, TB) + 1;
        w.Pnorm = eta + ld;
        w.PS = w.Pnorm + nt;
        w.Tv = w.PS + nt * 2 * NBMAX;
        w.Prow = w.Tv + 2 * NBMAX;
        w.ldp = ld;
        w.Pcol = w.Prow + (cdiv(nt, SEG) + 1) * ld;
        // strictly lower triangle of the input into the slot
        const double* src = a.A + a.offs[b];
        for (int64_t c = 0; c < n; c++)
            for (int64_t r = c + 1 + tid; r < n; r += WG) M[r + c * ld] = src[r + c * n];
        __syncthreads();
        tridiag_team1(M, ld, n, a.nb, w, tau, eta, sh, sm);
        logpf_reduce(eta, tau, n, a.logpf + b, a.sgn + b);
        if (a.fac) {
            double* dst = a.fac + a.offs[b];
            for (int64_t c = 0; c < n; c++)
                for (int64_t r = c + 1 + tid; r < n; r += WG) dst[r + c * n] = M[r + c * ld];
            double* te = a.taueta + a.toffs[b];
            for (int64_t i = tid; i < n - 1; i += WG) {
                te[i] = tau[i];
                te[n + i] = eta[i];
            }
            __syncthreads();
        }
    }
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
__global__ void k_larft(const double* Gm, int64_t ldg, const double* tau, int nbp, double* T, int64_t ldt)
{
    const int i = threadIdx.x;
    for (int k = 0; k < nbp; k++) {
        if (i < nbp) {
            double v = 0.0;
            if (i < k) {
                double acc = 0.0;
                for (int l = i; l < k; l++) acc += T[i + l * ldt] * Gm[l + k * ldg];
                v = -tau[k] * acc;
            } else if (i == k) {
                v = tau[k];
            }
            T[i + k * ldt] = v;
        }
        __syncthreads();
    }
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

__global__ void k_larft_batched(const double* Gm, int64_t gstride, const double* taueta, const int64_t* toffs,
                                int64_t c0, int nbp, double* T, int64_t tstride)
{
    const int z = blockIdx.x;
    const double* g = Gm + (int64_t)z * gstride;
    const double* tau = taueta + toffs[z] + c0;
    double* t = T + (int64_t)z * tstride;
    const int i = threadIdx.x;
    for (int k = 0; k < nbp; k++) {
        if (i < nbp) {
            double v = 0.0;
            if (i < k) {
                double acc = 0.0;
                for (int l = i; l < k; l++) acc += t[i + l * nbp] * g[l + k * nbp];
                v = -tau[k] * acc;
            } else if (i == k) {
                v = tau[k];
            }
            t[i + k * nbp] = v;
        }
        __syncthreads();
    }
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

extern "C" int skh_internal_launch_panel(const PanelLaunch* pl, hipStream_t st)
{
    void* args[] = {(void*)&pl->A, (void*)&pl->lda, (void*)&pl->n, (void*)&pl->c0, (void*)&pl->nbp, (void*)&pl->w,
                    (void*)&pl->tau, (void*)&pl->eta, (void*)&pl->bar, (void*)&pl->base, (void*)&pl->err};
    return (int)hipLaunchCooperativeKernel((const void*)k_panel, dim3(pl->grid), dim3(WG), args, 0, st);
}

extern "C" int skh_internal_panel_occupancy(int* blocks_per_cu)
{
    return (int)hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, (const void*)k_panel, WG, 0);
}

extern "C" int skh_internal_launch_rank2k(double* A, int64_t lda, const double* U, int64_t ldu, int64_t n, int64_t rlo,
                                          int64_t clo, int64_t c0, int nbp, hipStream_t st)
{
    TileGrid tg = tile_grid(n, rlo, clo);
    if (tg.total <= 0) return 0;
    int64_t per = (tg.total + 7) / 8;
    hipLaunchKernelGGL(k_rank2k, dim3((unsigned)(per * 8)), dim3(WG), 0, st, A, lda, U, ldu, n, rlo, clo, c0, nbp, tg,
                       per);
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

extern "C" int skh_internal_launch_batched(const BatchArgs* a, int grid, hipStream_t st)
{
    hipLaunchKernelGGL(k_batched, dim3(grid), dim3(WG), 0, st, *a);
    return (int)hipGetLastError();
}

extern "C" int skh_internal_batched_occupancy(int* blocks_per_cu)
{
    return (int)hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, (const void*)k_batched, WG, 0);
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
    hipLaunchKernelGGL(k_larft, dim3(1), dim3(128), 0, st, Gm, ldg, tau, nbp, T, ldt);
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
    hipLaunchKernelGGL(k_larft_batched, dim3((unsigned)count), dim3(128), 0, st, Gm, gstride, taueta, toffs, c0, nbp, T,
                       tstride);
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
