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

#include "skh_device.hpp"

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

// ------------------------------------------------------------------------------------------
// One large matrix over the whole GPU (cooperative launch, G workgroups).  The upper triangle, in
// position space (rows and columns as permuted so far), is cut into 64 x 64 tiles with fixed
// owners: tile k of the list ordered by decreasing block row goes to workgroup k mod G, so the
// tiles still inside the shrinking trailing square are always a prefix of the list and spread
// evenly.  A workgroup only ever reads and writes its own tiles with plain loads and stores; what
// crosses workgroups goes through agent-scope coherent accesses (ld_co / st_co) around the
// fence-free grid barrier: the argmax candidates, the norm partials and, each step, the "lines"
// of the (at most four) positions the step touches: i0, i1 and the pivot's row and column.  The
// line of position m holds the raw upper-storage entries U[min(t, m), max(t, m)] for every t.
// The step's swaps (ii <-> i0, then kk <-> i1) form a permutation s of those positions; the new
// entry at (a, b) is the old one at (s(a), s(b)) (negated when s reverses the order, for the
// trailing square; unchanged for the finished rows of R above i0).  Per step: candidates -> pivot
// -> publish lines -> barrier -> the new rows n0, n1 and the pivot block from the lines, the
// rank-2 update of the own tiles (old values of entries in moved rows or columns from the lines),
// the moved columns of R, new candidates -> barrier.  The arithmetic is that of chol_one.
// ------------------------------------------------------------------------------------------
constexpr int TB = 64;

struct CoopArgs {
    const double* A;   // n x n, ld n (read above the diagonal)
    double* U;         // n x n, ld n: R on return
    int64_t* p;        // n, 0-based
    int64_t* steps;    // 1
    double* W;         // 4 lines of n
    double* part;      // 3 G words: candidate value, packed position, norm partial
    unsigned* bar;     // grid barrier counters (zeroed before the launch)
    int* err;
    int n;
    int pivot;
};

__device__ __forceinline__ void tile_of(int64_t k, int nt, int& I, int& J)
{
    // k = m (m + 1) / 2 + (J - I) with m = nt - 1 - I
    int m = (int)((sqrt(8.0 * (double)k + 1.0) - 1.0) * 0.5);
    while ((int64_t)(m + 1) * (m + 2) / 2 <= k) m++;
    while ((int64_t)m * (m + 1) / 2 > k) m--;
    I = nt - 1 - m;
    J = I + (int)(k - (int64_t)m * (m + 1) / 2);
}

// packed candidate position (row, column) in a double-sized word
__device__ __forceinline__ double pack_pos(int r, int c)
{
    return __builtin_bit_cast(double, ((unsigned long long)(unsigned)c << 32) | (unsigned)r);
}
__device__ __forceinline__ void unpack_pos(double w, int& r, int& c)
{
    const unsigned long long u = __builtin_bit_cast(unsigned long long, w);
    r = (int)(u & 0xffffffffu);
    c = (int)(u >> 32);
}

constexpr int MAXT = 4;   // tiles whose loads a workgroup keeps in flight together

struct CoopShared {
    double nr[MAXT][2 * TB][2];   // new rows (n0, n1) at each tile's row range, then column range
    double bv[4];
    int br[4], bc[4];
    double red[4];
    double best, tol;
    int ii, jj;
    int ok;
};

// workgroup-wide min-order reduction of candidates into s.best / s.ii / s.jj (all threads call)
__device__ void coop_reduce(CoopShared& s, double bv, int br, int bc)
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

__global__ __launch_bounds__(CW) void k_skewchol_coop(CoopArgs g)
{
#pragma clang fp contract(off)
    __shared__ CoopShared s;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int n = g.n, G = gridDim.x, gi = blockIdx.x;
    const int nt = (n + TB - 1) / TB;
    const int64_t T = (int64_t)nt * (nt + 1) / 2;
    double* __restrict__ U = g.U;
    unsigned epoch = 0;

    // ---- copy the own tiles (strict upper part; zeros elsewhere), norm and first candidates ----
    double ss = 0.0, bv = 0.0;
    int br = 0, bc = 0;
    for (int64_t k = gi; k < T; k += G) {
        int I, J;
        tile_of(k, nt, I, J);
        const int a = I * TB + lane;
        for (int cc = wave; cc < TB; cc += CW / 64) {
            const int b = J * TB + cc;
            if (a < n && b < n) {
                const double v = a < b ? g.A[a + (int64_t)b * n] : 0.0;
                U[a + (int64_t)b * n] = v;
                ss += v * v;
                if (g.pivot && a < b) offer(bv, br, bc, v, a, b);
            }
        }
    }
    // blocks below the diagonal belong to no tile: zero them (never read in this launch)
    for (int64_t k = gi; k < (int64_t)nt * nt; k += G) {
        const int I = (int)(k % nt), J = (int)(k / nt);
        if (I <= J) continue;
        const int a = I * TB + lane;
        for (int cc = wave; cc < TB; cc += CW / 64) {
            const int b = J * TB + cc;
            if (a < n && b < n) U[a + (int64_t)b * n] = 0.0;
        }
    }
    if (gi == 0)
        for (int i = tid; i < n; i += CW) g.p[i] = i;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) ss += __shfl_xor(ss, o, 64);
    if (lane == 0) s.red[wave] = ss;
    coop_reduce(s, bv, br, bc);   // its barriers also publish s.red
    if (tid == 0) {
        st_co(g.part + 3 * gi, s.best);
        st_co(g.part + 3 * gi + 1, pack_pos(s.ii, s.jj));
        st_co(g.part + 3 * gi + 2, (s.red[0] + s.red[1]) + (s.red[2] + s.red[3]));
    }
    if (!grid_sync_co(g.bar, ++epoch, g.err, &s.ok)) return;
    // tol = 1e-15 ||A||_F, the same sum order in every workgroup
    {
        double t = 0.0;
        for (int q = tid; q < G; q += CW) t += ld_co(g.part + 3 * q + 2);
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) t += __shfl_xor(t, o, 64);
        if (lane == 0) s.red[wave] = t;
        __syncthreads();
        if (tid == 0) s.tol = 1e-15 * sqrt(2.0 * ((s.red[0] + s.red[1]) + (s.red[2] + s.red[3])));
        __syncthreads();
    }

    int64_t steps = 0;
    for (int j = 0; j < n / 2; j++) {
        const int i0 = 2 * j, i1 = 2 * j + 1, t0 = i1 + 1;
        // ---- pivot from every workgroup's candidate ----
        int ii = i0, kk = i1;
        if (g.pivot) {
            double v = 0.0;
            int r = i0, c = i0;
            for (int q = tid; q < G; q += CW) {
                int r2, c2;
                const double v2 = ld_co(g.part + 3 * q);
                unpack_pos(ld_co(g.part + 3 * q + 1), r2, c2);
                if (cand_better(v2, r2, c2, v, r, c)) {
                    v = v2;
                    r = r2;
                    c = c2;
                }
            }
            coop_reduce(s, v, r, c);
            if (fabs(s.best) < s.tol) break;   // the same words in every workgroup
            ii = s.ii;
            const int jj = s.jj;
            kk = jj == i0 ? ii : jj;
        }
        // the step's positions: slots 0..3 hold i0, i1, ii, kk (duplicates resolve to the first)
        const int P[4] = {i0, i1, ii, kk};
        auto slot = [&](int m) { return m == i0 ? 0 : m == i1 ? 1 : m == ii ? 2 : 3; };
        auto inP = [&](int m) { return m == i0 || m == i1 || m == ii || m == kk; };
        // s(a): the old position of new position a (kk <-> i1 first, then ii <-> i0)
        auto sig = [&](int a) {
            const int t = a == kk ? i1 : a == i1 ? kk : a;
            return t == ii ? i0 : t == i0 ? ii : t;
        };
        // ---- publish the lines of P from the own tiles ----
        for (int64_t k = gi; k < T; k += G) {
            int I, J;
            tile_of(k, nt, I, J);
            for (int e = 0; e < 4; e++) {
                const int m = P[e];
                if (slot(m) != e) continue;
                double* Wm = g.W + (int64_t)e * n;
                if (m / TB == J)   // column m: rows t <= m of this tile
                    for (int t = I * TB + tid; t < min(I * TB + TB, n); t += CW)
                        if (t <= m) st_co(Wm + t, U[t + (int64_t)m * n]);
                if (m / TB == I)   // row m: columns t >= m of this tile
                    for (int t = J * TB + tid; t < min(J * TB + TB, n); t += CW)
                        if (t >= m) st_co(Wm + t, U[m + (int64_t)t * n]);
            }
        }
        if (!grid_sync_co(g.bar, ++epoch, g.err, &s.ok)) return;
        // the permutation (workgroup 0 alone; off the path to the barrier above)
        if (gi == 0 && tid == 0) {
            if (ii != i0) {
                const int64_t t = g.p[ii];
                g.p[ii] = g.p[i0];
                g.p[i0] = t;
            }
            if (kk != i1) {
                const int64_t t = g.p[kk];
                g.p[kk] = g.p[i1];
                g.p[i1] = t;
            }
        }
        // new trailing entry (x, y), x != y, both >= i0, with x or y moved: the old one at
        // (s(x), s(y)), whose line is published
        auto lineval = [&](int u, int v) {
            return inP(u) ? ld_co(g.W + (int64_t)slot(u) * n + v) : ld_co(g.W + (int64_t)slot(v) * n + u);
        };
        auto snew = [&](int x, int y) {
            const int u = sig(x), v = sig(y);
            const double w = lineval(u, v);
            return u < v ? w : -w;
        };
        // pivot block
        const double pv = snew(i0, i1);
        double sg = 1.0, r;
        if (g.pivot) {
            r = sqrt(pv);
        } else {
            sg = pv > 0.0 ? 1.0 : pv < 0.0 ? -1.0 : pv;
            r = sqrt(fabs(pv));
        }
        const double q = -1.0 / r;
        const bool moved = ii != i0 || kk != i1;
        // ---- own tiles ----
        bv = 0.0;
        br = t0;
        bc = t0;
        const int I0 = i0 / TB;   // block row of the pivot rows (i0 and i1 share it: TB is even)
        const int64_t K = (int64_t)(nt - I0) * (nt - I0 + 1) / 2;   // tiles at or below block row I0
        // finished rows of R in moved columns: U[a, b] = old U[a, s(b)], a < i0 (any own tile)
        if (moved)
            for (int64_t k = gi; k < T; k += G) {
                int I, J;
                tile_of(k, nt, I, J);
                const int r0 = I * TB;
                for (int e = 0; e < 4; e++) {
                    const int b = P[e];
                    if (slot(b) != e || b / TB != J || sig(b) == b) continue;
                    const int sb = sig(b);
                    for (int a = r0 + tid; a < min(r0 + TB, i0); a += CW)
                        U[a + (int64_t)b * n] = ld_co(g.W + (int64_t)slot(sb) * n + a);
                }
            }
        // the own tiles at or below the pivot rows, MAXT at a time: first the new rows at every
        // tile's row and column ranges (all line reads in flight together), then every entry's old
        // value (all tile loads in flight together), then the update
        for (int64_t kc = gi; kc < K; kc += (int64_t)MAXT * G) {
            int nc = 0;
            int tr0[MAXT], tc0[MAXT];
            for (int u = 0; u < MAXT; u++) {
                const int64_t k = kc + (int64_t)u * G;
                int I = 0, J = 0;
                if (k < K) {
                    tile_of(k, nt, I, J);
                    nc = u + 1;
                }
                tr0[u] = I * TB;
                tc0[u] = J * TB;
            }
            for (int t = tid; t < nc * 2 * TB; t += CW) {
                const int u = t / (2 * TB), v = t % (2 * TB);
                int r0u = tr0[0], c0u = tc0[0];   // constant indices only (no scratch)
#pragma unroll
                for (int x = 1; x < MAXT; x++)
                    if (u == x) {
                        r0u = tr0[x];
                        c0u = tc0[x];
                    }
                const int c = (v < TB ? r0u : c0u) + (v & (TB - 1));
                double x0 = 0.0, x1 = 0.0;
                if (c > i1 && c < n) {
                    const double b0 = snew(i0, c), b1 = snew(i1, c);
                    x0 = (g.pivot ? b1 : sg * b1) * q;
                    x1 = (-b0) * q;
                }
                s.nr[u][v][0] = x0;
                s.nr[u][v][1] = x1;
            }
            __syncthreads();
            double old[MAXT][TB / (CW / 64)];
#pragma unroll
            for (int u = 0; u < MAXT; u++) {
                const int a = tr0[u] + lane;
                const bool amov = moved && a != sig(a);
#pragma unroll
                for (int w = 0; w < TB / (CW / 64); w++) {
                    const int b = tc0[u] + wave + w * (CW / 64);
                    double o = 0.0;
                    if (u < nc && a < n && b < n && a >= t0 && a <= b) {
                        const bool mv = amov || (moved && b != sig(b));
                        o = !mv ? U[a + (int64_t)b * n] : a == b ? lineval(sig(a), sig(a)) : snew(a, b);
                    }
                    old[u][w] = o;
                }
            }
#pragma unroll
            for (int u = 0; u < MAXT; u++) {
                if (u >= nc) break;
                const int a = tr0[u] + lane;
#pragma unroll
                for (int w = 0; w < TB / (CW / 64); w++) {
                    const int cc = wave + w * (CW / 64);
                    const int b = tc0[u] + cc;
                    if (b >= n || a >= n || a > b) continue;
                    double* __restrict__ e = U + a + (int64_t)b * n;
                    if (a >= t0) {
                        const double x = s.nr[u][lane][0] * s.nr[u][TB + cc][1];
                        const double y = s.nr[u][lane][1] * s.nr[u][TB + cc][0];
                        const double wv = old[u][w] - (x - y);
                        *e = wv;
                        if (g.pivot && (a < b || isnan(wv))) offer(bv, br, bc, wv, a, b);
                    } else if (a == i0 || a == i1) {
                        if (b > i1) *e = a == i0 ? s.nr[u][TB + cc][0] : s.nr[u][TB + cc][1];
                        else if (b == i1 && a == i0) *e = 0.0;
                        else if (b == a) *e = (a == i0 || g.pivot) ? r : sg * r;
                    }
                }
            }
            __syncthreads();
        }
        steps++;
        if (g.pivot && j + 1 < n / 2) {
            coop_reduce(s, bv, br, bc);
            if (tid == 0) {
                st_co(g.part + 3 * gi, s.best);
                st_co(g.part + 3 * gi + 1, pack_pos(s.ii, s.jj));
            }
        }
        if (!grid_sync_co(g.bar, ++epoch, g.err, &s.ok)) return;
    }
    if (gi == 0 && tid == 0) *g.steps = steps;
}

}  // namespace chol
}  // namespace skh

using namespace skh;

extern "C" int skh_internal_skewchol_coop(const double* A, double* U, int64_t n, int pivot, int64_t* p, int64_t* steps,
                                          double* W, double* part, unsigned* bar, int* err, int grid, hipStream_t st)
{
    chol::CoopArgs a;
    a.A = A;
    a.U = U;
    a.p = p;
    a.steps = steps;
    a.W = W;
    a.part = part;
    a.bar = bar;
    a.err = err;
    a.n = (int)n;
    a.pivot = pivot;
    void* args[] = {(void*)&a};
    return (int)hipLaunchCooperativeKernel((const void*)chol::k_skewchol_coop, dim3(grid), dim3(chol::CW), args, 0, st);
}

extern "C" int skh_internal_skewchol_coop_occupancy(int* per_cu)
{
    return (int)hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, (const void*)chol::k_skewchol_coop, chol::CW, 0);
}

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
