// skh_compact.hpp -- compact engine: one workgroup per matrix, for batches of small and
// medium matrices (n <= NMAX), persistent over a work queue.
//
// Per column k the engine runs the one-reflector Householder step of the reference's skew
// reduction (householder!, src/hessenberg.jl:26-44; the update A22 += v w^T - w v^T with
// w = tau A22 v, src/hessenberg.jl:158-166 for a panel of one column):
//
//   (1) every thread: w_{k-1}[r] = tau_{k-1} y_{k-1}[r] for its rows; column k of the updated
//       matrix x_r = A[r,k] + v_{k-1}[r] w_{k-1}[k] - w_{k-1}[r] v_{k-1}[k]; sum of squares;
//       barrier; reflector k from (x_{k+1}, sum x_r^2 for r >= k+2), computed redundantly by
//       every thread from the same LDS values (bit-identical); v_k into LDS and into A[:,k];
//       barrier;
//   (2) one fused pass over the strictly lower triangle of the trailing block [k+1, n)^2:
//       a_rc += v_{k-1}[r] w_{k-1}[c] - w_{k-1}[r] v_{k-1}[c] (stored back), and the skew
//       matvec y_k = A22 v_k from the updated a_rc still in registers:
//       yr[r] += a_rc v_k[c], yc[c] += a_rc v_k[r], y_k = yr - yc;  barrier.
//
// Each element of the trailing triangle is read and written once per column; the pass has no
// barrier inside: row sums are flushed per strip into a per-wave LDS row vector owned by one
// lane, column sums into an LDS column vector owned by one lane (fixed ownership, so every
// reduction has a fixed order and the results are deterministic).
//
// Memory: matrix b column-major with ld = n (n even), 16-byte aligned; the factorization runs
// in a per-workgroup slot (logpf only) or in place in the caller's factor buffer (gradient).
#pragma once
#include "skh_device.hpp"

namespace skh {
namespace cmp {

constexpr int CW = 256;   // threads per workgroup (4 waves)

template <int NMAX>
struct CShared {
    double v[2][NMAX];   // v_k and v_{k-1} (ping-pong)
    double w[NMAX];      // w_{k-1}
    double yr[4][NMAX];  // per-wave row sums of the matvec
    double yc[NMAX];     // column sums of the matvec
    double red[4];
    double alpha;
    int64_t idx;
    int stop;
};

// One 64x64 tile (strip I, tile column J) of the trailing block: this lane's 4 rows x 4 columns
// (wave -> 16 columns, lane>>4 -> 4 of them, lane&15 -> rows {2p, 2p+1, 32+2p, 33+2p}).
struct TileRegs {
    double a[4][4];
};

// Every lane loads, entries outside the triangle from the matrix's first two elements (which always
// exist): a load under a per-lane branch, or a select right after it, is waited for on the spot,
// which would serialise the tile's eight loads.  The garbage in those entries is never stored (the
// store is masked) and never reaches a sum (the matvec masks again).
__device__ __forceinline__ void tile_load(const double* __restrict__ M, int64_t n, int64_t s0, int64_t I, int64_t J,
                                          int wave, int cg, int p, TileRegs& t)
{
    const int64_t cb = J * TB + 16 * wave + 4 * cg;
#pragma unroll
    for (int cc = 0; cc < 4; cc++) {
        const int64_t c = cb + cc;
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int64_t r = I * TB + 32 * h + 2 * p;   // even; r + 1 < n whenever r < n (n even)
            const bool ok = c >= s0 && c < n && r < n && r + 1 > c;
            const double2 v2 = *reinterpret_cast<const double2*>(M + (ok ? c * n + r : 0));
            t.a[2 * h][cc] = v2.x;
            t.a[2 * h + 1][cc] = v2.y;
        }
    }
}

__device__ __forceinline__ void tile_store(double* __restrict__ M, int64_t n, int64_t s0, int64_t I, int64_t J,
                                           int wave, int cg, int p, const TileRegs& t)
{
    const int64_t cb = J * TB + 16 * wave + 4 * cg;
#pragma unroll
    for (int cc = 0; cc < 4; cc++) {
        const int64_t c = cb + cc;
        double* col = M + c * n;
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int64_t r = I * TB + 32 * h + 2 * p;
            if (c >= s0 && c < n && r < n) {
                if (r > c) {
                    double2 v2;
                    v2.x = t.a[2 * h][cc];
                    v2.y = t.a[2 * h + 1][cc];
                    *reinterpret_cast<double2*>(col + r) = v2;
                } else if (r + 1 > c) {
                    col[r + 1] = t.a[2 * h + 1][cc];
                }
            }
        }
    }
}

// The fused update + matvec pass over the strictly lower triangle of [s0, n)^2.
template <int NMAX>
__device__ void fused_pass(double* __restrict__ M, int64_t n, int64_t s0, bool upd, const double* __restrict__ vp,
                           const double* __restrict__ wp, const double* __restrict__ vn, CShared<NMAX>& s)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int p = lane & 15, cg = lane >> 4;
    const int64_t nt = cdiv(n, TB);
    const int64_t I0 = s0 / TB;
    int64_t I = I0, J = I0;
    TileRegs cur, nxt;
    tile_load(M, n, s0, I, J, wave, cg, p, cur);
    double rv[4], rw[4], rn[4], racc[4];
#pragma unroll
    for (int rr = 0; rr < 4; rr++) {
        const int64_t r = I * TB + 32 * (rr >> 1) + 2 * p + (rr & 1);
        rv[rr] = vp[r];
        rw[rr] = wp[r];
        rn[rr] = vn[r];
        racc[rr] = 0.0;
    }
    for (;;) {
        const bool last_in_strip = (J == I);
        const int64_t In = last_in_strip ? I + 1 : I;
        const int64_t Jn = last_in_strip ? I0 : J + 1;
        const bool more = In < nt;
        if (more) tile_load(M, n, s0, In, Jn, wave, cg, p, nxt);
        const int64_t cb = J * TB + 16 * wave + 4 * cg;
        double cv[4], cw[4], cn[4], cacc[4];
#pragma unroll
        for (int cc = 0; cc < 4; cc++) {
            cv[cc] = vp[cb + cc];
            cw[cc] = wp[cb + cc];
            cn[cc] = vn[cb + cc];
            cacc[cc] = 0.0;
        }
        if (upd) {
#pragma unroll
            for (int rr = 0; rr < 4; rr++)
#pragma unroll
                for (int cc = 0; cc < 4; cc++) cur.a[rr][cc] += rv[rr] * cw[cc] - rw[rr] * cv[cc];
            tile_store(M, n, s0, I, J, wave, cg, p, cur);
        }
        // entries outside the triangle hold whatever was loaded for them: mask
#pragma unroll
        for (int cc = 0; cc < 4; cc++) {
            const int64_t c = cb + cc;
#pragma unroll
            for (int rr = 0; rr < 4; rr++) {
                const int64_t r = I * TB + 32 * (rr >> 1) + 2 * p + (rr & 1);
                const double a = (c >= s0 && r > c && r < n) ? cur.a[rr][cc] : 0.0;
                racc[rr] += a * cn[cc];
                cacc[cc] += a * rn[rr];
            }
        }
#pragma unroll
        for (int cc = 0; cc < 4; cc++) {
            const double x = row16_sum(cacc[cc]);
            if (p == 0) s.yc[cb + cc] += x;   // column cb+cc is always owned by this lane
        }
        if (last_in_strip) {
#pragma unroll
            for (int rr = 0; rr < 4; rr++) {
                const double x = xor32_sum(xor16_sum(racc[rr]));
                if (cg == 0) s.yr[wave][I * TB + 32 * (rr >> 1) + 2 * p + (rr & 1)] = x;
                racc[rr] = 0.0;
            }
            if (more) {
#pragma unroll
                for (int rr = 0; rr < 4; rr++) {
                    const int64_t r = In * TB + 32 * (rr >> 1) + 2 * p + (rr & 1);
                    rv[rr] = vp[r];
                    rw[rr] = wp[r];
                    rn[rr] = vn[r];
                }
            }
        }
        if (!more) break;
        cur = nxt;
        I = In;
        J = Jn;
    }
}

// Watchdog: workgroup-uniform stop flag once the launch has run longer than `budget` ticks of
// the 100 MHz wall clock (0 = no limit), or once another workgroup has raised the error flag.
template <int NMAX>
__device__ bool should_stop(CShared<NMAX>& s, int* err, unsigned long long t_start, unsigned long long budget)
{
    if (threadIdx.x == 0) {
        int st = 0;
        if (budget != 0 && wall_clock64() - t_start > budget) {
            atomicOr(err, 2);
            st = 1;
        }
        if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) st = 1;
        s.stop = st;
    }
    __syncthreads();
    const bool r = s.stop != 0;
    __syncthreads();
    return r;
}

// Tridiagonalize one matrix M (n even, 2 <= n <= NMAX, ld = n) in place: factors in the strictly
// lower triangle ('L' layout, unit stored), tau[0..n-2], eta[0..n-2].  Returns false if stopped.
template <int NMAX>
__device__ bool tridiag_one(double* __restrict__ M, int64_t n, double* tau, double* eta, CShared<NMAX>& s, int* err,
                            unsigned long long t_start, unsigned long long budget)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    constexpr int RPT = NMAX / CW;   // rows per thread (row r = tid + CW * i)
    for (int r = tid; r < NMAX; r += CW) {
        s.v[0][r] = 0.0;
        s.v[1][r] = 0.0;
        s.w[r] = 0.0;
        s.yc[r] = 0.0;
        s.yr[0][r] = 0.0;
        s.yr[1][r] = 0.0;
        s.yr[2][r] = 0.0;
        s.yr[3][r] = 0.0;
    }
    __syncthreads();
    int cur = 0;
    double tprev = 0.0;
    for (int64_t k = 0; k <= n - 2; k++) {
        if ((k & 31) == 31 && should_stop(s, err, t_start, budget)) return false;
        const bool upd = k >= 1;
        const double* vp = s.v[cur ^ 1];
        double wk = 0.0;
        if (upd) wk = tprev * (((s.yr[0][k] + s.yr[1][k]) + (s.yr[2][k] + s.yr[3][k])) - s.yc[k]);
        const double vpk = vp[k];
        double x[RPT];
        double sq = 0.0;
#pragma unroll
        for (int i = 0; i < RPT; i++) {
            const int64_t r = tid + CW * i;
            x[i] = 0.0;
            if (r >= k && r < n) {
                double wr = 0.0;
                if (upd) {
                    wr = tprev * (((s.yr[0][r] + s.yr[1][r]) + (s.yr[2][r] + s.yr[3][r])) - s.yc[r]);
                    s.w[r] = wr;
                }
                if (r > k) {
                    // row k is read by every thread (wk) in this phase and never touched again
                    s.yr[0][r] = 0.0;
                    s.yr[1][r] = 0.0;
                    s.yr[2][r] = 0.0;
                    s.yr[3][r] = 0.0;
                    s.yc[r] = 0.0;
                    double xv = M[r + k * n];
                    if (upd) xv += vp[r] * wk - wr * vpk;
                    x[i] = xv;
                    if (r == k + 1) s.alpha = xv;
                    else sq += xv * xv;
                }
            }
        }
        sq = wave_sum(sq);
        if (lane == 0) s.red[wave] = sq;
        __syncthreads();
        const double alpha = s.alpha;
        if (k == n - 2) {
            // last 1x1 block (real length-1 reflector: tau = 0, beta = x)
            if (tid == 0) {
                eta[k] = alpha;
                tau[k] = 0.0;
                M[(k + 1) + k * n] = alpha;
            }
            break;
        }
        const ColScal cs = reflector_scalars(alpha, (s.red[0] + s.red[1]) + (s.red[2] + s.red[3]));
        double* vn = s.v[cur];
#pragma unroll
        for (int i = 0; i < RPT; i++) {
            const int64_t r = tid + CW * i;
            double v = 0.0;
            if (r > k && r < n) {
                v = (r == k + 1) ? cs.unit : cs.scale * x[i];
                M[r + k * n] = v;
            }
            vn[r] = v;
        }
        if (tid == 0) {
            tau[k] = cs.tau;
            eta[k] = cs.beta;
        }
        __syncthreads();
        fused_pass<NMAX>(M, n, k + 1, upd, vp, s.w, vn, s);
        __syncthreads();
        tprev = cs.tau;
        cur ^= 1;
    }
    __syncthreads();
    return true;
}

}  // namespace cmp

}  // namespace skh
