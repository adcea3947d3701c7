// skh_grad.hip -- the backward pass of batched log|Pf|: d log|Pf(A)| / dA = -1/2 A^{-1}.
//
//   A = Q T Q^T (the forward factorization), A^{-1} = Q T^{-1} Q^T,
//   T^{-1} = P^T [[0, -C^{-T}], [C^{-1}, 0]] P with C (n/2 x n/2, lower bidiagonal)
//   C[a,a] = -eta[2a], C[a,a-1] = eta[2a-1];  Q_O = Q[:, 0::2], Q_E = Q[:, 1::2],
//   K = Q_E C^{-1} Q_O^T,  A^{-1} = K - K^T.
//
// Per size group of matrices (grid.z = member of the group, all of size n):
//   Q = H_0 ... H_{n-2} from the stored reflectors, blocked compact WY from the last block of
//   32 reflectors to the first: Q22 <- Q22 - V (T (V^T Q22)), T = larft('F', 'C') of V^T V;
//   Zt = (C^{-1} Q_O^T)^T by forward substitution; K = Q_E Zt^T; G = -1/2 (K - K^T).
// Every product runs in k_bgemm, a batched FP64 GEMM on v_mfma_f64_16x16x4f64.  The factors
// and the gradient share the caller's buffer (factors in, gradient out): member z of a group
// lives at base + goff[z]; its tau | eta at taueta + toff[z].  Workspaces are strided.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "skh_internal.h"

namespace skh {
namespace grad {

typedef double v4d __attribute__((ext_vector_type(4)));

// ---- batched GEMM --------------------------------------------------------------------------
// C_z = alpha op(A_z) op(B_z) + beta C_z, column-major, X_z = X + z * sX.  A workgroup of four
// waves owns a 64 x 64 tile of C, each wave a 32 x 32 quarter (2 x 2 blocks of 16 x 16); the
// operands stream from L2 into the MFMA registers one k-step (4) ahead, no LDS, no barrier.
// v_mfma_f64_16x16x4f64 lanes (gfx950): a-operand lane l = (row l%16, k l/16), b-operand lane
// l = (k l/16, col l%16), accumulator item q = (row 4q + l/16, col l%16).  Each block is formed
// transposed, D = op(B)_blk^T op(A)_blk^T, so item q of lane l is C[r + l%16, c + 4q + l/16]
// and a wave's stores are 128-byte column segments.  beta == 0 does not read C.
template <bool TA, bool TB>
__global__ __launch_bounds__(256) void k_bgemm(int M, int N, int K, double alpha, const double* __restrict__ A,
                                               int64_t lda, int64_t sA, const double* __restrict__ B, int64_t ldb,
                                               int64_t sB, double beta, double* __restrict__ C, int64_t ldc, int64_t sC)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, lc = lane & 15, lq = lane >> 4;
    const int64_t z = blockIdx.z;
    A += z * sA;
    B += z * sB;
    C += z * sC;
    const int r0 = blockIdx.x * 64 + 32 * (w & 1);
    const int c0 = blockIdx.y * 64 + 32 * (w >> 1);
    if (r0 >= M || c0 >= N) return;   // the whole wave lies outside C (no barriers here)
    int ra[2], cb[2];
    bool va[2], vb[2];
#pragma unroll
    for (int b = 0; b < 2; b++) {
        ra[b] = r0 + 16 * b + lc;
        va[b] = ra[b] < M;
        ra[b] = va[b] ? ra[b] : M - 1;
        cb[b] = c0 + 16 * b + lc;
        vb[b] = cb[b] < N;
        cb[b] = vb[b] ? cb[b] : N - 1;
    }
    auto load = [&](int k0, double af[2], double bf[2]) {
        const int k = k0 + lq;
        const bool kok = k < K;
        const int64_t kc = kok ? k : 0;
#pragma unroll
        for (int b = 0; b < 2; b++) {
            const double x = TA ? A[kc + (int64_t)ra[b] * lda] : A[ra[b] + kc * lda];
            af[b] = (kok && va[b]) ? x : 0.0;
            const double y = TB ? B[cb[b] + kc * ldb] : B[kc + (int64_t)cb[b] * ldb];
            bf[b] = (kok && vb[b]) ? y : 0.0;
        }
    };
    v4d acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = v4d{0.0, 0.0, 0.0, 0.0};
    double a0[2], b0[2], a1[2], b1[2];
    if (K > 0) load(0, a0, b0);
    for (int k0 = 0; k0 < K; k0 += 8) {
        const bool two = k0 + 4 < K;
        if (two) load(k0 + 4, a1, b1);
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(b0[j], a0[i], acc[i][j], 0, 0, 0);
        if (!two) break;
        if (k0 + 8 < K) load(k0 + 8, a0, b0);
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(b1[j], a1[i], acc[i][j], 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 2; i++) {
        const int r = r0 + 16 * i + lc;
        if (r >= M) continue;
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int c = c0 + 16 * j + 4 * q + lq;
                if (c < N) {
                    double* p = C + r + (int64_t)c * ldc;
                    const double v = alpha * acc[i][j][q];
                    *p = (beta != 0.0) ? v + beta * *p : v;
                }
            }
    }
}

// ---- helpers of the backward pass ---------------------------------------------------------
// Q_z = I (n x n, ld n)
__global__ void k_eye_s(double* Q, int64_t qstride, int64_t n)
{
    double* q = Q + (int64_t)blockIdx.z * qstride;
    const int64_t c = blockIdx.y;
    for (int64_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x)
        q[r + c * n] = (r == c) ? 1.0 : 0.0;
}

// V_z (m x nbp, ld m) = reflectors c0 .. c0+nbp-1 of member z (unit on the head, zeros above)
__global__ void k_form_v_o(const double* F, const int64_t* goff, int64_t n, int64_t c0, int nbp, double* V,
                           int64_t vstride)
{
    const int k = blockIdx.y;
    const double* f = F + goff[blockIdx.z];
    double* v = V + (int64_t)blockIdx.z * vstride;
    const int64_t m = n - c0 - 1;
    for (int64_t rr = blockIdx.x * blockDim.x + threadIdx.x; rr < m; rr += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = c0 + 1 + rr, head = c0 + k + 1;
        v[rr + k * m] = (r == head) ? 1.0 : ((r > head) ? f[r + (c0 + k) * n] : 0.0);
    }
}

// Zt_z (n x n/2, ld n) = (C^{-1} Q_O^T)^T: one thread per row c of Q_z, forward substitution along a
__global__ void k_bidiag_solve_s(const double* Q, int64_t qstride, const double* taueta, const int64_t* toff,
                                 int64_t n, double* Zt, int64_t zstride)
{
    const int z = blockIdx.y;
    const double* q = Q + (int64_t)z * qstride;
    const double* eta = taueta + toff[z] + n;
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

// G_z = alpha (K_z - K_z^T) through 32 x 33 LDS tiles, G_z at G + goff[z] (ld n)
__global__ void k_skew_out_o(const double* K, int64_t kstride, double* G, const int64_t* goff, int64_t n,
                             double alpha)
{
    __shared__ double tile[32][33];
    const double* k = K + (int64_t)blockIdx.z * kstride;
    double* g = G + goff[blockIdx.z];
    const int64_t r0 = (int64_t)blockIdx.x * 32, c0 = (int64_t)blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8 threads
    for (int yy = ty; yy < 32; yy += 8) {
        const int64_t r = c0 + tx, c = r0 + yy;   // the K^T block: K[c0 + tx, r0 + yy]
        tile[yy][tx] = (r < n && c < n) ? k[r + c * n] : 0.0;
    }
    __syncthreads();
    for (int yy = ty; yy < 32; yy += 8) {
        const int64_t r = r0 + tx, c = c0 + yy;
        if (r < n && c < n) g[r + c * n] = alpha * (k[r + c * n] - tile[tx][yy]);
    }
}

// zero the n x n block of every listed member (odd n: log|Pf| = -Inf and a zero gradient)
__global__ void k_zero_members(double* G, const int64_t* goff, const int64_t* gn)
{
    double* g = G + goff[blockIdx.y];
    const int64_t cnt = gn[blockIdx.y] * gn[blockIdx.y];
    for (int64_t i = blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += (int64_t)gridDim.x * blockDim.x) g[i] = 0.0;
}

}  // namespace grad
}  // namespace skh

using namespace skh::grad;

extern "C" int skh_internal_bgemm(int ta, int tb, int64_t M, int64_t N, int64_t K, double alpha, const double* A,
                                  int64_t lda, int64_t sA, const double* B, int64_t ldb, int64_t sB, double beta,
                                  double* C, int64_t ldc, int64_t sC, int count, hipStream_t st)
{
    if (M <= 0 || N <= 0 || count <= 0) return 0;
    if (M > (1 << 30) || N > (1 << 30) || K > (1 << 30)) return (int)hipErrorInvalidValue;
    const dim3 grid((unsigned)((M + 63) / 64), (unsigned)((N + 63) / 64), (unsigned)count);
    const int m = (int)M, n = (int)N, k = (int)K;
    if (!ta && !tb)
        hipLaunchKernelGGL((k_bgemm<false, false>), grid, dim3(256), 0, st, m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C,
                           ldc, sC);
    else if (ta && !tb)
        hipLaunchKernelGGL((k_bgemm<true, false>), grid, dim3(256), 0, st, m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C,
                           ldc, sC);
    else if (!ta && tb)
        hipLaunchKernelGGL((k_bgemm<false, true>), grid, dim3(256), 0, st, m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C,
                           ldc, sC);
    else
        hipLaunchKernelGGL((k_bgemm<true, true>), grid, dim3(256), 0, st, m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C,
                           ldc, sC);
    return (int)hipGetLastError();
}

extern "C" int skh_internal_grad_eye(double* Q, int64_t qstride, int64_t n, int count, hipStream_t st)
{
    hipLaunchKernelGGL(k_eye_s, dim3((unsigned)std::min<int64_t>((n + 255) / 256, 16), (unsigned)n, (unsigned)count),
                       dim3(256), 0, st, Q, qstride, n);
    return (int)hipGetLastError();
}

extern "C" int skh_internal_grad_form_v(const double* F, const int64_t* goff, int64_t n, int64_t c0, int nbp, double* V,
                                        int64_t vstride, int count, hipStream_t st)
{
    const int64_t m = n - c0 - 1;
    hipLaunchKernelGGL(k_form_v_o, dim3((unsigned)std::min<int64_t>((m + 255) / 256, 16), (unsigned)nbp, (unsigned)count),
                       dim3(256), 0, st, F, goff, n, c0, nbp, V, vstride);
    return (int)hipGetLastError();
}

extern "C" int skh_internal_grad_bidiag(const double* Q, int64_t qstride, const double* taueta, const int64_t* toff,
                                        int64_t n, double* Zt, int64_t zstride, int count, hipStream_t st)
{
    hipLaunchKernelGGL(k_bidiag_solve_s, dim3((unsigned)std::max<int64_t>(1, (n + 63) / 64), (unsigned)count), dim3(64),
                       0, st, Q, qstride, taueta, toff, n, Zt, zstride);
    return (int)hipGetLastError();
}

extern "C" int skh_internal_grad_skew_out(const double* K, int64_t kstride, double* G, const int64_t* goff, int64_t n,
                                          double alpha, int count, hipStream_t st)
{
    const unsigned t = (unsigned)((n + 31) / 32);
    hipLaunchKernelGGL(k_skew_out_o, dim3(t, t, (unsigned)count), dim3(256), 0, st, K, kstride, G, goff, n, alpha);
    return (int)hipGetLastError();
}

extern "C" int skh_internal_grad_zero(double* G, const int64_t* goff, const int64_t* gn, int count, hipStream_t st)
{
    if (count <= 0) return 0;
    hipLaunchKernelGGL(k_zero_members, dim3(64, (unsigned)count), dim3(256), 0, st, G, goff, gn);
    return (int)hipGetLastError();
}
