// skh_chol.hip -- skew-Cholesky factorization R^T J R = A[p, p] of real skew-symmetric matrices on
// gfx950 (P. Benner et al., "Cholesky-like factorizations of skew-symmetric matrices", ETNA 11,
// 2000), one workgroup per matrix of a batch.
//
// Reference: _skewchol!(A::SkewHermitian{<:Real}, ::RowMaximum) and (..., ::NoPivot) in
// src/cholesky.jl:58-107 and :28-56.  Step j (pivot rows i0 = 2j, i1 = 2j + 1): with complete
// pivoting the largest entry of the trailing square (findmax: first in column-major order among
// equal values) moves to (i0, i1) by symmetric swaps of rows and columns; a maximum below
// tol = 1e-15 ||A||_F ends the factorization.  Then, with r = sqrt(A[i0, i1]), q = -1/r and the old
// rows b0 = A[i0, c], b1 = A[i1, c] (c > i1), the new rows of R are n0 = b1 q and n1 = (-b0) q, the
// diagonal is (r, r), and the trailing square receives A[a, b] -= n0[a] n1[b] - n1[a] n0[b].
// Without pivoting r = sqrt|A[i0, i1]|, s = sign(A[i0, i1]), R[i1, i1] = s r and n0 = (s b1) q.
//
// Storage: the factorization runs in the output R (n x n, ld = n, column-major).  The trailing
// square is kept as its upper triangle: the update's two products are rounded separately (no fused
// multiply-add), so the increment is exactly antisymmetric and the square stays exactly
// skew-symmetric, which makes the lower half redundant (its diagonal stays 0 unless an Inf or NaN
// reaches it).  The pivot search of the next step is fused into the update: each updated entry
// offers the positive one of its two mirror images.  Traffic per matrix: the trailing triangle
// read and written once per step, 2 x 8 B x n^3 / 12 in all.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

namespace skh {
namespace chol {

constexpr int CW = 256;          // threads per workgroup
constexpr int NSMALL = 256;      // LDS variants: the two new rows take 2 x 8 B x NM (4 or 64 KiB)
constexpr int NMAX = 4096;       // largest n

template <int NM>
struct CShared {
    double n0[NM];
    double n1[NM];
    double bv[4];
    int br[4], bc[4];
    double best;
    int ii, jj;
    double tol;
};

// Julia's findmax order on (value, position): NaN above every number; among equal values the
// earlier position in column-major order (smaller column, then smaller row).
__device__ __forceinline__ bool cand_better(double v2, int r2, int c2, double v, int r, int c)
{
    const bool n1 = isnan(v), n2 = isnan(v2);
    if (n1 != n2) return n2;
    if (n2 || v2 == v) return c2 < c || (c2 == c && r2 < r);
    return v2 > v;
}

// The upper entry (a, b), a < b, of the skew square with value v: v sits at (a, b) and -v at
// (b, a).  Offers the positive one; a NaN at its earlier position (b, a).  Zeros never beat the
// default candidate, the square's first diagonal entry.
__device__ __forceinline__ void offer(double& bv, int& br, int& bc, double v, int a, int b)
{
    double w;
    int r, c;
    if (v > 0.0) {
        w = v;
        r = a;
        c = b;
    } else if (v < 0.0) {
        w = -v;
        r = b;
        c = a;
    } else if (isnan(v)) {
        w = v;
        r = b;
        c = a;
    } else {
        return;
    }
    if (cand_better(w, r, c, bv, br, bc)) {
        bv = w;
        br = r;
        bc = c;
    }
}

// workgroup-wide best candidate into s.best / s.ii / s.jj (every thread must call)
template <int NM>
__device__ void reduce_best(CShared<NM>& s, double bv, int br, int bc)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        const double v2 = __shfl_xor(bv, o, 64);
        const int r2 = __shfl_xor(br, o, 64), c2 = __shfl_xor(bc, o, 64);
        if (cand_better(v2, r2, c2, bv, br, bc)) {
            bv = v2;
            br = r2;
            bc = c2;
        }
    }
    if (lane == 0) {
        s.bv[wave] = bv;
        s.br[wave] = br;
        s.bc[wave] = bc;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double v = s.bv[0];
        int r = s.br[0], c = s.bc[0];
        for (int w = 1; w < CW / 64; w++)
            if (cand_better(s.bv[w], s.br[w], s.bc[w], v, r, c)) {
                v = s.bv[w];
                r = s.br[w];
                c = s.bc[w];
            }
        s.best = v;
        s.ii = r;
        s.jj = c;
    }
    __syncthreads();
}

// Symmetric swap of indices x < y in upper storage (U[a, b], a < b; U[b, a] = -U[a, b]).  Rows
// a < x include the finished rows of R, whose columns x and y trade places as in the reference's
// whole-column swap (:77-79).
__device__ void swap_sym(double* __restrict__ U, int n, int x, int y)
{
    const int tid = threadIdx.x;
    for (int a = tid; a < x; a += CW) {
        const double t = U[a + (int64_t)x * n];
        U[a + (int64_t)x * n] = U[a + (int64_t)y * n];
        U[a + (int64_t)y * n] = t;
    }
    for (int b = x + 1 + tid; b < y; b += CW) {
        const double t1 = U[x + (int64_t)b * n], t2 = U[b + (int64_t)y * n];
        U[x + (int64_t)b * n] = -t2;
        U[b + (int64_t)y * n] = -t1;
    }
    if (tid == 0) {
        U[x + (int64_t)y * n] = -U[x + (int64_t)y * n];
        const double t = U[x + (int64_t)x * n];   // diagonal entries (0 unless an Inf or NaN reached them)
        U[x + (int64_t)x * n] = U[y + (int64_t)y * n];
        U[y + (int64_t)y * n] = t;
    }
    for (int b = y + 1 + tid; b < n; b += CW) {
        const double t = U[x + (int64_t)b * n];
        U[x + (int64_t)b * n] = U[y + (int64_t)b * n];
        U[y + (int64_t)b * n] = t;
    }
}

// One matrix: A (ld n, read above the diagonal only) -> R (ld n), p (0-based), *steps.
template <int NM>
__device__ void chol_one(const double* __restrict__ A, double* __restrict__ U, int n, int pivot, CShared<NM>& s,
                         int64_t* __restrict__ p, int64_t* steps_out)
{
#pragma clang fp contract(off)
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // copy the strict upper triangle, zero the rest; tol from the Frobenius norm of the full square
    double ss = 0.0;
    double bv = 0.0;
    int br = 0, bc = 0;
    for (int c = 0; c < n; c++)
        for (int r = tid; r < n; r += CW) {
            const double v = r < c ? A[r + (int64_t)c * n] : 0.0;
            U[r + (int64_t)c * n] = v;
            ss += v * v;
            if (pivot && r < c) offer(bv, br, bc, v, r, c);
        }
    for (int i = tid; i < n; i += CW) p[i] = i;
    if (pivot) {
        // sum of squares over the workgroup (reusing the candidate slots' LDS after a barrier)
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) ss += __shfl_xor(ss, o, 64);
        __syncthreads();
        if (lane == 0) s.n0[wave] = ss;
        __syncthreads();
        if (tid == 0) {
            double t = 0.0;
            for (int w = 0; w < CW / 64; w++) t += s.n0[w];
            s.tol = 1e-15 * sqrt(2.0 * t);
        }
        reduce_best(s, bv, br, bc);   // ends with a barrier, so s.tol is visible
    } else {
        __syncthreads();
    }
    int64_t steps = 0;
    for (int j = 0; j < n / 2; j++) {
        const int i0 = 2 * j, i1 = 2 * j + 1;
        double sg = 1.0;
        if (pivot) {
            if (fabs(s.best) < s.tol) break;   // uniform: every thread reads the same LDS words
            const int ii = s.ii, jj = s.jj;
            const int kk = jj == i0 ? ii : jj;
            if (ii != i0) {
                swap_sym(U, n, min(ii, i0), max(ii, i0));
                if (tid == 0) {
                    const int64_t t = p[ii];
                    p[ii] = p[i0];
                    p[i0] = t;
                }
                __syncthreads();
            }
            if (kk != i1) {
                swap_sym(U, n, min(kk, i1), max(kk, i1));
                if (tid == 0) {
                    const int64_t t = p[kk];
                    p[kk] = p[i1];
                    p[i1] = t;
                }
                __syncthreads();
            }
        }
        const double v = U[i0 + (int64_t)i1 * n];
        double r;
        if (pivot) {
            r = sqrt(v);
        } else {
            sg = v > 0.0 ? 1.0 : v < 0.0 ? -1.0 : v;   // Julia's sign: +-0 and NaN pass through
            r = sqrt(fabs(v));
        }
        const double q = -1.0 / r;
        for (int c = i1 + 1 + tid; c < n; c += CW) {
            const double b0 = U[i0 + (int64_t)c * n], b1 = U[i1 + (int64_t)c * n];
            const double x0 = (pivot ? b1 : sg * b1) * q;
            const double x1 = (-b0) * q;
            s.n0[c] = x0;
            s.n1[c] = x1;
            U[i0 + (int64_t)c * n] = x0;
            U[i1 + (int64_t)c * n] = x1;
        }
        if (tid == 0) {
            U[i0 + (int64_t)i0 * n] = r;
            U[i1 + (int64_t)i1 * n] = pivot ? r : sg * r;
            U[i0 + (int64_t)i1 * n] = 0.0;
        }
        __syncthreads();
        // trailing update U[a, b] -= n0[a] n1[b] - n1[a] n0[b], i1 < a <= b (the diagonal's increment
        // is 0 unless an Inf or NaN is involved, as in the reference's full-square gemm); wave w takes
        // columns i1 + 1 + w, + 4, ...; the next pivot's candidates on the way
        const int t0 = i1 + 1;
        bv = 0.0;
        br = t0;
        bc = t0;
        for (int b = t0 + wave; b < n; b += CW / 64) {
            const double nb0 = s.n0[b], nb1 = s.n1[b];
            double* __restrict__ col = U + (int64_t)b * n;
            for (int a = t0 + lane; a <= b; a += 64) {
                const double x = s.n0[a] * nb1, y = s.n1[a] * nb0;
                const double w = col[a] - (x - y);
                col[a] = w;
                if (pivot && (a < b || isnan(w))) offer(bv, br, bc, w, a, b);
            }
        }
        steps++;
        if (pivot && j + 1 < n / 2) reduce_best(s, bv, br, bc);
        else __syncthreads();
    }
    if (tid == 0 && steps_out) *steps_out = steps;
    __syncthreads();
}

// One workgroup per matrix; order[] lists the matrices (largest first), workgroup g takes items
// g, g + G, ...  Matrix b: A + offs[b] and R + offs[b] (ld n[b]), p + poffs[b].
template <int NM>
__global__ __launch_bounds__(CW) void k_skewchol(const double* __restrict__ A, double* __restrict__ R,
                                                const int64_t* __restrict__ ns, const int64_t* __restrict__ offs,
                                                const int64_t* __restrict__ poffs, const int64_t* __restrict__ order,
                                                int64_t batch, int pivot, int64_t* __restrict__ p,
                                                int64_t* __restrict__ steps)
{
    __shared__ CShared<NM> s;
    for (int64_t q = blockIdx.x; q < batch; q += gridDim.x) {
        const int64_t b = order[q];
        chol_one<NM>(A + offs[b], R + offs[b], (int)ns[b], pivot, s, p + poffs[b], steps ? steps + b : nullptr);
    }
}

}  // namespace chol
}  // namespace skh

using namespace skh;

extern "C" int skh_internal_skewchol(const double* A, double* R, const int64_t* ns, const int64_t* offs,
                                     const int64_t* poffs, const int64_t* order, int64_t batch, int64_t nmax,
                                     int pivot, int64_t* p, int64_t* steps, int grid, hipStream_t st)
{
    if (batch <= 0) return 0;
    if (nmax <= chol::NSMALL)
        hipLaunchKernelGGL(chol::k_skewchol<chol::NSMALL>, dim3((unsigned)grid), dim3(chol::CW), 0, st, A, R, ns, offs,
                           poffs, order, batch, pivot, p, steps);
    else
        hipLaunchKernelGGL(chol::k_skewchol<chol::NMAX>, dim3((unsigned)grid), dim3(chol::CW), 0, st, A, R, ns, offs,
                           poffs, order, batch, pivot, p, steps);
    return (int)hipGetLastError();
}

extern "C" int skh_internal_skewchol_occupancy(int64_t nmax, int* per_cu)
{
    const void* f = nmax <= chol::NSMALL ? (const void*)chol::k_skewchol<chol::NSMALL>
                                         : (const void*)chol::k_skewchol<chol::NMAX>;
    return (int)hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, f, chol::CW, 0);
}

extern "C" int skh_internal_skewchol_nmax() { return chol::NMAX; }
