// skh_cblock.hpp -- blocked engine for batches: one workgroup per matrix (n <= NMAX), panels of NB
// reflectors, the blocked reduction of SkewLinearAlgebra.jl (src/hessenberg.jl:92-174) inside one
// workgroup.
//
// Per panel [c0, c0 + NB) and column j = c0 + i of it:
//   X  x = A[:, j] + V W[j, :]^T - W V[j, :]^T on the rows r > j (V, W: the panel's reflectors and
//      w vectors so far, k < i), the norm of x[j+2:];                                  barrier
//   S  the reflector of x (householder!, src/hessenberg.jl:26-44); v into A[:, j]; the partial sums
//      of t1 = W^T v and t2 = V^T v;                                                   barrier
//   M  the skew matvec y = A22 v over the strictly lower triangle of the trailing block
//      [j+1, n)^2 of the matrix as it was at the start of the panel (read only: the panel's
//      updates enter through V and W), 64 x 64 tiles, the next tile in flight;        barrier
//   W  w = tau (y + V t1 - W t2) on the rows r > j;                                    barrier
// and after the panel the skew rank-2k update A22 += V W^T - W V^T of the trailing lower triangle
// [c0 + NB, n)^2 on v_mfma_f64_16x16x4f64 (one wave per 32 x 32 tile, operands from L2).
// Thread t owns the rows r = t + CW u (u < NMAX / CW) and keeps their V and W entries of the panel
// in registers; row j of V and W goes to LDS for everyone's column update.  Each trailing element
// is read once per column and read and written once per panel, against once each per column in
// the unblocked engine, and the per-element work of a column is two FMAs instead of four.
#pragma once
#include "skh_device.hpp"

namespace skh {
namespace cbk {

constexpr int CW = 256;   // threads per workgroup (4 waves)

// column-sum partials of the matvec: one per quad of lane rows (after a quad reduction), each
// address written by one lane, so the sums are deterministic
template <int NMAX>
struct ColParts {
    static constexpr int PC = 4;
};

template <int NMAX, int NB>
struct BShared {
    double v[NMAX];          // v of the current column (0 outside (j, n))
    double yr[4][NMAX];      // per-wave row sums of the matvec
    double ycp[ColParts<NMAX>::PC][NMAX];   // column-sum partials of the matvec
    double vj[NB], wj[NB];   // row j of the panel's V and W (k < i)
    double tred[4][2 * NB];  // per-wave partials of t1 = W^T v (first NB) and t2 = V^T v
    double red[4];
    double alpha;
    int64_t idx;
    int stop;
};

// One lane's share of a 64 x 64 tile (strip I, column block J) of the trailing block: rows
// I*64 + {2p, 2p+1, 32+2p, 33+2p}, columns J*64 + 16 wave + 4 cg + cc.  Entries outside the
// strictly lower triangle of [s0, n)^2 read as zero (n even, so r + 1 < n whenever r < n).
struct Tile {
    double a[4][4];
};

// 32-bit index math: n <= 1024, so every element offset is below 2^20.  Every lane loads (entries
// outside the triangle from the matrix's first two elements, which always exist) and the mask is
// applied when the tile is used, so the eight loads of a tile are issued back to back and waited
// for one iteration later (a load under a per-lane branch, or a select right after it, is waited
// for on the spot).
__device__ __forceinline__ bool tile_ok(int n, int s0, int c, int r) { return c >= s0 && c < n && r < n && r + 1 > c; }

__device__ __forceinline__ void tile_ld(const double* __restrict__ M, int n, int s0, int I, int J, int wave, int cg,
                                        int p, Tile& t)
{
    const int cb = J * TB + 16 * wave + 4 * cg;
#pragma unroll
    for (int cc = 0; cc < 4; cc++) {
        const int c = cb + cc;
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int r = I * TB + 32 * h + 2 * p;
            const double2 v2 = *reinterpret_cast<const double2*>(M + (tile_ok(n, s0, c, r) ? c * n + r : 0));
            t.a[2 * h][cc] = v2.x;
            t.a[2 * h + 1][cc] = v2.y;
        }
    }
}

// zero the entries of a loaded tile outside the strictly lower triangle of [s0, n)^2
__device__ __forceinline__ void tile_mask(int n, int s0, int I, int J, int wave, int cg, int p, Tile& t)
{
    const int cb = J * TB + 16 * wave + 4 * cg;
#pragma unroll
    for (int cc = 0; cc < 4; cc++) {
        const int c = cb + cc;
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int r = I * TB + 32 * h + 2 * p;
            const bool ok = tile_ok(n, s0, c, r);
            t.a[2 * h][cc] = (ok && r > c) ? t.a[2 * h][cc] : 0.0;
            t.a[2 * h + 1][cc] = ok ? t.a[2 * h + 1][cc] : 0.0;
        }
    }
}

// the next tile in strip-major order over the lower triangle of tile blocks [I0, nt)^2
__device__ __forceinline__ void tile_next(int I0, int& I, int& J)
{
    if (J == I) {
        I++;
        J = I0;
    } else {
        J++;
    }
}

// y = A22 v (row sums into yr[wave], column sums into ycp) over the strictly lower triangle of
// [s0, n)^2; `cur` holds the first tile (I0, I0), loaded (not masked) by the caller.
template <int NMAX, int NB>
__device__ void matvec_ro(const double* __restrict__ M, int n, int s0, Tile& cur, BShared<NMAX, NB>& s)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int p = lane & 15, cg = lane >> 4;
    const int nt = (n + TB - 1) / TB;
    const int I0 = s0 / TB;
    int I = I0, J = I0;
    double rn[4], racc[4];
#pragma unroll
    for (int rr = 0; rr < 4; rr++) {
        rn[rr] = s.v[I * TB + 32 * (rr >> 1) + 2 * p + (rr & 1)];
        racc[rr] = 0.0;
    }
    for (;;) {
        int In = I, Jn = J;
        tile_next(I0, In, Jn);
        const bool more = In < nt;
        Tile nxt;
        if (more) tile_ld(M, n, s0, In, Jn, wave, cg, p, nxt);
        tile_mask(n, s0, I, J, wave, cg, p, cur);
        const int cb = J * TB + 16 * wave + 4 * cg;
        double cv[4], cacc[4];
#pragma unroll
        for (int cc = 0; cc < 4; cc++) {
            cv[cc] = s.v[cb + cc];
            cacc[cc] = 0.0;
        }
#pragma unroll
        for (int cc = 0; cc < 4; cc++)
#pragma unroll
            for (int rr = 0; rr < 4; rr++) {
                racc[rr] += cur.a[rr][cc] * cv[cc];
                cacc[cc] += cur.a[rr][cc] * rn[rr];
            }
        // column partials: added into LDS without a return (no wait); the address (partial, column)
        // belongs to one lane in every strip, so the additions keep a fixed order
        constexpr int PC = ColParts<NMAX>::PC;
#pragma unroll
        for (int cc = 0; cc < 4; cc++) {
            double x = cacc[cc];
            if (PC == 4) {
                x += dpp_d<DPP_XOR1>(x);
                x += dpp_d<DPP_XOR2>(x);
            }
            if (PC == 16 || (p & 3) == 0)
                __hip_atomic_fetch_add(&s.ycp[PC == 16 ? p : (p >> 2)][cb + cc], x, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_WORKGROUP);
        }
        if (J == I) {   // last tile of strip I: its row sums
#pragma unroll
            for (int rr = 0; rr < 4; rr++) {
                const double x = xor32_sum(xor16_sum(racc[rr]));
                if (cg == 0) s.yr[wave][I * TB + 32 * (rr >> 1) + 2 * p + (rr & 1)] = x;
                racc[rr] = 0.0;
            }
            if (more) {
#pragma unroll
                for (int rr = 0; rr < 4; rr++) rn[rr] = s.v[In * TB + 32 * (rr >> 1) + 2 * p + (rr & 1)];
            }
        }
        if (!more) break;
        cur = nxt;
        I = In;
        J = Jn;
    }
}

// Skew rank-2k update of one 32 x 32 tile of the trailing lower triangle (rows r0.., columns q0..):
// A[r, c] += sum_k V[r,k] W[c,k] - W[r,k] V[c,k] for r > c >= e, r < n; V[:, k] = A[:, c0 + k],
// W[:, k] = Wg[:, k] (ld n), k < nbp.  L = [V W], R = [W -V]; operand rows past n are clamped (their
// products land in rows or columns that are not stored).  Lane layout as in the other MFMA kernels:
// item q of lane l of block (bi, bj) is C[r0 + 16 bi + l%16, q0 + 16 bj + 4q + l/16].
template <int NB>
__device__ void update_tile(double* __restrict__ M, int64_t n, const double* __restrict__ Wg, int64_t c0, int nbp,
                            int64_t e, int64_t r0, int64_t q0)
{
    const int lane = threadIdx.x & 63, lc = lane & 15, lq = lane >> 4;
    int64_t rl[2], rq[2];
#pragma unroll
    for (int b = 0; b < 2; b++) {
        rl[b] = min(r0 + 16 * b + lc, n - 1);
        rq[b] = min(q0 + 16 * b + lc, n - 1);
    }
    v4d acc[2][2];
#pragma unroll
    for (int bi = 0; bi < 2; bi++)
#pragma unroll
        for (int bj = 0; bj < 2; bj++) {
            const int64_t row = r0 + 16 * bi + lc;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int64_t col = q0 + 16 * bj + 4 * q + lq;
                const bool ok = row < n && col >= e && col < row;
                const double c = M[ok ? row + col * n : 0];   // unconditional load, masked after
                acc[bi][bj][q] = ok ? c : 0.0;
            }
        }
    // every operand of the tile is loaded before the first product (one memory round trip)
    constexpr int KS = NB / 4;   // k-steps per half
    double lv[KS][2], lw[KS][2], rv[KS][2], rw[KS][2];
#pragma unroll
    for (int s = 0; s < KS; s++) {
        const int kk = s * 4 + lq;
        const int64_t kc = kk < nbp ? kk : 0;   // in-bounds address for the lanes that read nothing
#pragma unroll
        for (int b = 0; b < 2; b++) {
            lv[s][b] = M[rl[b] + (c0 + kc) * n];
            lw[s][b] = Wg[rl[b] + kc * n];
            rw[s][b] = Wg[rq[b] + kc * n];
            rv[s][b] = M[rq[b] + (c0 + kc) * n];
        }
    }
#pragma unroll
    for (int s = 0; s < 2 * KS; s++) {
        const int h = s / KS, ks = s % KS;
        const bool ok = ks * 4 + lq < nbp;
        double lf[2], rf[2];
#pragma unroll
        for (int b = 0; b < 2; b++) {
            lf[b] = ok ? (h == 0 ? lv[ks][b] : lw[ks][b]) : 0.0;
            rf[b] = ok ? (h == 0 ? rw[ks][b] : -rv[ks][b]) : 0.0;
        }
#pragma unroll
        for (int bi = 0; bi < 2; bi++)
#pragma unroll
            for (int bj = 0; bj < 2; bj++)
                acc[bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(rf[bj], lf[bi], acc[bi][bj], 0, 0, 0);
    }
#pragma unroll
    for (int bi = 0; bi < 2; bi++) {
        const int64_t row = r0 + 16 * bi + lc;
#pragma unroll
        for (int bj = 0; bj < 2; bj++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int64_t col = q0 + 16 * bj + 4 * q + lq;
                if (row < n && col >= e && col < row) M[row + col * n] = acc[bi][bj][q];
            }
    }
}

// Tridiagonalize one matrix M (n even, 2 <= n <= NMAX, ld n) in place, W workspace Wg (NMAX x NB):
// factors in the strictly lower triangle ('L' layout, unit stored), tau[0..n-2], eta[0..n-2].
// Returns false if stopped by the watchdog.
template <int NMAX, int NB, typename StopFn>
__device__ bool tridiag_blocked(double* __restrict__ M, int64_t n, double* __restrict__ Wg, double* tau, double* eta,
                                BShared<NMAX, NB>& s, StopFn stop)
{
    constexpr int RPT = NMAX / CW;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int p = lane & 15, cg = lane >> 4;
    constexpr int PC = ColParts<NMAX>::PC;
    for (int r = tid; r < NMAX; r += CW) {
        s.v[r] = 0.0;
#pragma unroll
        for (int q = 0; q < PC; q++) s.ycp[q][r] = 0.0;
#pragma unroll
        for (int q = 0; q < 4; q++) s.yr[q][r] = 0.0;
    }
    __syncthreads();
    double xa[RPT];
#pragma unroll
    for (int u = 0; u < RPT; u++) {
        const int64_t r = tid + CW * u;
        xa[u] = (r > 0 && r < n) ? M[r] : 0.0;   // column 0
    }
    for (int64_t c0 = 0; c0 < n - 1; c0 += NB) {
        if (stop()) return false;
        const int nbp = (int)min<int64_t>(NB, n - 1 - c0);
        const int64_t e = c0 + nbp;   // first column after the panel
        double vr[RPT][NB], wr[RPT][NB];
#pragma unroll
        for (int u = 0; u < RPT; u++)
#pragma unroll
            for (int k = 0; k < NB; k++) {
                vr[u][k] = 0.0;
                wr[u][k] = 0.0;
            }
        for (int i = 0; i < nbp; i++) {
            const int64_t j = c0 + i;
            // the matvec's first tile does not depend on v: in flight through X and S
            Tile first;
            const bool mv = j + 2 < n;
            if (mv) tile_ld(M, (int)n, (int)(j + 1), (int)((j + 1) / TB), (int)((j + 1) / TB), wave, cg, p, first);
            // ---- X: column j of the panel-updated matrix, its norm ----
            double x[RPT];
            double sq = 0.0;
#pragma unroll
            for (int u = 0; u < RPT; u++) {
                const int64_t r = tid + CW * u;
                x[u] = 0.0;
                if (r > j && r < n) {
                    double xv = xa[u];
#pragma unroll
                    for (int k = 0; k < NB; k++)
                        if (k < i) xv += vr[u][k] * s.wj[k] - wr[u][k] * s.vj[k];
                    x[u] = xv;
                    if (r == j + 1) s.alpha = xv;
                    else sq += xv * xv;
                }
            }
            sq = wave_sum(sq);
            if (lane == 0) s.red[wave] = sq;
            __syncthreads();
            // ---- S: reflector, v, partial t ----
            const double alpha = s.alpha;
            if (j == n - 2) {
                // last 1x1 block (real length-1 reflector: tau = 0, beta = x)
                if (tid == 0) {
                    eta[j] = alpha;
                    tau[j] = 0.0;
                    M[(j + 1) + j * n] = alpha;
                }
                __syncthreads();
                return true;
            }
            const ColScal cs = reflector_scalars(alpha, (s.red[0] + s.red[1]) + (s.red[2] + s.red[3]));
            double v[RPT];
#pragma unroll
            for (int u = 0; u < RPT; u++) {
                const int64_t r = tid + CW * u;
                double vv = 0.0;
                if (r > j && r < n) {
                    vv = (r == j + 1) ? cs.unit : cs.scale * x[u];
                    M[r + j * n] = vv;
                }
                v[u] = vv;
                s.v[r] = vv;
#pragma unroll
                for (int k = 0; k < NB; k++)
                    if (k == i) vr[u][k] = vv;
            }
            if (tid == 0) {
                tau[j] = cs.tau;
                eta[j] = cs.beta;
            }
            // per-wave partials of t1 = W^T v and t2 = V^T v (k < i; zero above)
#pragma unroll
            for (int k = 0; k < NB; k++) {
                double a1 = 0.0, a2 = 0.0;
                if (k < i) {   // workgroup-uniform
#pragma unroll
                    for (int u = 0; u < RPT; u++) {
                        a1 += wr[u][k] * v[u];
                        a2 += vr[u][k] * v[u];
                    }
                    a1 = wave_sum(a1);
                    a2 = wave_sum(a2);
                }
                if (lane == 0) {
                    s.tred[wave][k] = a1;
                    s.tred[wave][NB + k] = a2;
                }
            }
            // the next column of the panel (unchanged until the panel's update)
#pragma unroll
            for (int u = 0; u < RPT; u++) {
                const int64_t r = tid + CW * u;
                xa[u] = (j + 1 < e && r > j + 1 && r < n) ? M[r + (j + 1) * n] : 0.0;
            }
            __syncthreads();
            // ---- M: y = A22 v ----
            if (mv) matvec_ro<NMAX, NB>(M, (int)n, (int)(j + 1), first, s);
            __syncthreads();
            // ---- W: w = tau (y + V t1 - W t2) ----
#pragma unroll
            for (int u = 0; u < RPT; u++) {
                const int64_t r = tid + CW * u;
                if (r > j && r < n) {
                    double yc = 0.0;
#pragma unroll
                    for (int q = 0; q < PC; q++) {
                        yc += s.ycp[q][r];
                        s.ycp[q][r] = 0.0;
                    }
                    const double y = ((s.yr[0][r] + s.yr[1][r]) + (s.yr[2][r] + s.yr[3][r])) - yc;
                    double acc = y;
#pragma unroll
                    for (int k = 0; k < NB; k++)
                        if (k < i) {
                            const double t1 = (s.tred[0][k] + s.tred[1][k]) + (s.tred[2][k] + s.tred[3][k]);
                            const double t2 = (s.tred[0][NB + k] + s.tred[1][NB + k]) + (s.tred[2][NB + k] + s.tred[3][NB + k]);
                            acc += vr[u][k] * t1 - wr[u][k] * t2;
                        }
                    const double wv = cs.tau * acc;
#pragma unroll
                    for (int k = 0; k < NB; k++)
                        if (k == i) wr[u][k] = wv;
#pragma unroll
                    for (int q = 0; q < 4; q++) s.yr[q][r] = 0.0;
                    if (r == j + 1) {   // row j + 1 of V and W for the next column's update
#pragma unroll
                        for (int k = 0; k < NB; k++) {
                            s.vj[k] = vr[u][k];
                            s.wj[k] = wr[u][k];
                        }
                    }
                }
            }
            __syncthreads();
        }
        // ---- the panel's skew rank-2k update of the trailing lower triangle [e, n)^2 ----
        if (e >= n - 1) continue;   // the next (last) column needs no update: nothing trails it
#pragma unroll
        for (int u = 0; u < RPT; u++) {
            const int64_t r = tid + CW * u;
            if (r >= e && r < n)
#pragma unroll
                for (int k = 0; k < NB; k++)
                    if (k < nbp) Wg[r + (int64_t)k * n] = wr[u][k];
        }
        __syncthreads();
        {
            const int64_t b0 = e / 32, nb32 = cdiv(n, 32) - b0;
            const int64_t tiles = nb32 * (nb32 + 1) / 2;
            for (int64_t t = wave; t < tiles; t += 4) {
                int64_t bi, bj;
                tri_ij(t, bi, bj);
                update_tile<NB>(M, n, Wg, c0, nbp, e, (b0 + bi) * 32, (b0 + bj) * 32);
            }
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < RPT; u++) {
            const int64_t r = tid + CW * u;
            xa[u] = (r > e && r < n) ? M[r + e * n] : 0.0;
        }
    }
    __syncthreads();
    return true;
}

}  // namespace cbk
}  // namespace skh
