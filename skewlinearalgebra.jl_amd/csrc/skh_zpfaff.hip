// skh_zpfaff.hip -- Pfaffians of complex skew-symmetric matrices (A^T = -A, ComplexF64) on
// gfx950: Parlett-Reid elimination with pivoting, one workgroup per matrix of a batch.
//
// Reference: _pfaffian!(A::AbstractMatrix{<:Complex}) and _logabspfaffian!(...) in
// src/pfaffian.jl:97-137 and :193-233.  Step k (0-based, even): the row r > k of column k with
// the largest modulus (first one on ties, as Julia's argmax) is swapped with row and column
// k + 1 (the Pfaffian changes sign); a zero A[k+1, k] ends with Pf = 0; otherwise
// Pf *= A[k, k+1] and, with tau_c = A[k, c] / A[k, k+1] and col_c = A[c, k+1] (c >= k+2), the
// trailing block receives A[r, c] += tau_r col_c - tau_c col_r.
//
// The reference updates the full square; this kernel keeps only the strictly lower triangle of
// a working copy (the update's increment is antisymmetric in (r, c), so an exactly
// skew-symmetric input stays exactly skew-symmetric and the upper half is implied:
// A[k, c] = -A[c, k]).  Traffic per matrix: the trailing triangle read and written once per step,
// 2 x 16 B x n^3 / 12 in all.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <algorithm>

#include "skh_device.hpp"

// Products and sums are rounded separately everywhere in this file (the complex helpers
// included), so the Pfaffian is bit-identical to the oracle's, which does the same arithmetic.
#pragma clang fp contract(off)

namespace skh {
namespace zpf {

constexpr int PW = 256;          // threads per workgroup
constexpr int NSMALL = 256;      // LDS variants: tau and col take 2 x 16 B x NM (8 KiB or 64 KiB)
constexpr int NMAX = 2048;       // largest n

struct cplx {
    double re, im;
};
__device__ __forceinline__ cplx cmul(cplx a, cplx b) { return cplx{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ cplx csub(cplx a, cplx b) { return cplx{a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ cplx cadd(cplx a, cplx b) { return cplx{a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ cplx cneg(cplx a) { return cplx{-a.re, -a.im}; }
// a / b (Smith's algorithm: scaled by the larger component of b)
__device__ __forceinline__ cplx cdiv(cplx a, cplx b)
{
    if (fabs(b.re) >= fabs(b.im)) {
        const double r = b.im / b.re, d = b.re + b.im * r;
        return cplx{(a.re + a.im * r) / d, (a.im - a.re * r) / d};
    }
    const double r = b.re / b.im, d = b.re * r + b.im;
    return cplx{(a.re * r + a.im) / d, (a.im * r - a.re) / d};
}
__device__ __forceinline__ cplx ld2(const double2* p) { const double2 t = *p; return cplx{t.x, t.y}; }
__device__ __forceinline__ void st2(double2* p, cplx v) { *p = make_double2(v.re, v.im); }

template <int NM>
struct PShared {
    cplx tau[NM];
    cplx col[NM];
    double bv[4];
    int bi[4];
    int kp;
};

// Keeps the (modulus, index) pair that Julia's argmax keeps: NaN above every number, the
// smaller index among equal keys.
__device__ __forceinline__ void amax_merge(double& v, int& i, double v2, int i2)
{
    const bool n1 = isnan(v), n2 = isnan(v2);
    const bool take = n1 != n2 ? n2 : (n2 || v2 == v) ? i2 < i : v2 > v;
    if (take) {
        v = v2;
        i = i2;
    }
}

// Parlett-Reid on the strictly lower triangle of L (n x n, ld = n, column-major).  mode 0: the
// Pfaffian into o1[0..1]; mode 1: log|Pf| into o1[0] and cis(phase) into o2[0..1].
template <int NM>
__device__ void pfaff_one(double2* __restrict__ L, int n, int mode, PShared<NM>& s, double* o1, double* o2)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    cplx pf = cplx{1.0, 0.0};
    double logpf = 0.0, phase = 0.0;
    bool zero = (n % 2) == 1;
    for (int k = 0; k + 1 < n && !zero; k += 2) {
        // ---- pivot: largest |L[r, k]|, r > k (the sentinel -1 loses to every row) ----
        double bv = -1.0;
        int bi = n;
        for (int r = k + 1 + tid; r < n; r += PW) {
            const cplx a = ld2(L + r + (int64_t)k * n);
            amax_merge(bv, bi, hypot(a.re, a.im), r);
        }
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) amax_merge(bv, bi, __shfl_xor(bv, o, 64), __shfl_xor(bi, o, 64));
        if (lane == 0) {
            s.bv[wave] = bv;
            s.bi[wave] = bi;
        }
        __syncthreads();
        if (tid == 0) {
            double v = s.bv[0];
            int i = s.bi[0];
            for (int w = 1; w < 4; w++) amax_merge(v, i, s.bv[w], s.bi[w]);
            s.kp = i;
        }
        __syncthreads();
        const int kp = s.kp;
        if (kp != k + 1) {
            // symmetric swap of k+1 and kp on the indices >= k, in lower storage (A[x,y] = -L[y,x])
            if (tid == 0) {
                const cplx a = ld2(L + (k + 1) + (int64_t)k * n), b = ld2(L + kp + (int64_t)k * n);
                st2(L + (k + 1) + (int64_t)k * n, b);
                st2(L + kp + (int64_t)k * n, a);
                st2(L + kp + (int64_t)(k + 1) * n, cneg(ld2(L + kp + (int64_t)(k + 1) * n)));
            }
            for (int x = kp + 1 + tid; x < n; x += PW) {   // rows below kp: columns k+1 <-> kp
                const cplx a = ld2(L + x + (int64_t)(k + 1) * n), b = ld2(L + x + (int64_t)kp * n);
                st2(L + x + (int64_t)(k + 1) * n, b);
                st2(L + x + (int64_t)kp * n, a);
            }
            for (int y = k + 2 + tid; y < kp; y += PW) {   // between: row kp <-> column k+1, negated
                const cplx a = ld2(L + kp + (int64_t)y * n), b = ld2(L + y + (int64_t)(k + 1) * n);
                st2(L + kp + (int64_t)y * n, cneg(b));
                st2(L + y + (int64_t)(k + 1) * n, cneg(a));
            }
            if (mode == 0) pf = cneg(pf);
            else phase += 3.14159265358979323846;   // Float64(pi)
            __syncthreads();
        }
        const cplx sub = ld2(L + (k + 1) + (int64_t)k * n);   // A[k+1, k]
        if (sub.re == 0.0 && sub.im == 0.0) {
            zero = true;
            break;
        }
        const cplx a = cneg(sub);   // A[k, k+1]
        if (mode == 0) pf = cmul(pf, a);
        else {
            logpf += log(hypot(a.re, a.im));
            phase += atan2(a.im, a.re);
        }
        // ---- tau_c = A[k, c] / A[k, k+1] = -L[c, k] / a, col_c = L[c, k+1], c >= k+2 ----
        for (int c = k + 2 + tid; c < n; c += PW) {
            s.tau[c] = cdiv(cneg(ld2(L + c + (int64_t)k * n)), a);
            s.col[c] = ld2(L + c + (int64_t)(k + 1) * n);
        }
        __syncthreads();
        // ---- L[r, c] += tau_r col_c - tau_c col_r, r > c >= k+2: wave w takes columns w, w+4, ... ----
        for (int c = k + 2 + wave; c < n - 1; c += 4) {
            const cplx tc = s.tau[c], cc = s.col[c];
            double2* Lc = L + (int64_t)c * n;
            for (int r = c + 1 + lane; r < n; r += 64) {
                const cplx d = csub(cmul(s.tau[r], cc), cmul(tc, s.col[r]));
                st2(Lc + r, cadd(ld2(Lc + r), d));
            }
        }
        __syncthreads();
    }
    if (tid == 0) {
        if (mode == 0) {
            o1[0] = zero ? 0.0 : pf.re;
            o1[1] = zero ? 0.0 : pf.im;
        } else {
            o1[0] = zero ? -INFINITY : logpf;
            o2[0] = zero ? 0.0 : cos(phase);
            o2[1] = zero ? 0.0 : sin(phase);
        }
    }
    __syncthreads();
}

// One workgroup per matrix (order[] lists the matrices, largest first; workgroup g takes items
// g, g + G, ...).  The strictly lower triangle of matrix b is copied into the workgroup's slot
// (nmax x nmax, ld = n[b]) and eliminated there; the input is not modified.
template <int NM>
__global__ __launch_bounds__(PW) void k_zpfaff(const double2* __restrict__ A, const int64_t* __restrict__ ns,
                                              const int64_t* __restrict__ offs, const int64_t* __restrict__ order,
                                              int64_t batch, double2* slots, int64_t slot, int mode, double* o1,
                                              double* o2)
{
    __shared__ PShared<NM> s;
    double2* L = slots + (int64_t)blockIdx.x * slot;
    for (int64_t q = blockIdx.x; q < batch; q += gridDim.x) {
        const int64_t b = order[q];
        const int n = (int)ns[b];
        const double2* M = A + offs[b];
        for (int64_t c = 0; c < n; c++)
            for (int r = (int)c + 1 + (int)threadIdx.x; r < n; r += PW) L[r + c * n] = M[r + c * n];
        __syncthreads();
        pfaff_one(L, n, mode, s, o1 + (mode == 0 ? 2 : 1) * b, o2 + 2 * b);
    }
}

// ------------------------------------------------------------------------------------------
// One large matrix over the whole GPU (cooperative launch, G workgroups).  The strictly lower
// triangle, in position space, is cut into 64 x 64 tiles with fixed owners: tile k of the list in
// decreasing block column goes to workgroup k mod G, so the tiles that still meet the shrinking
// trailing square are a prefix of the list.  A workgroup reads and writes only its own tiles
// (plain loads and stores); what crosses workgroups goes through agent-scope coherent accesses
// around the fence-free grid barrier: the pivot candidates of column k (offered by the owners of
// that column during the previous step's update) and, each step, the "lines" of positions k, k+1
// and the pivot row kp: the raw lower entries L[max(t, m), min(t, m)] for t >= k.  The swap
// k+1 <-> kp is the permutation s; the new entry (r, c) is the old one at (s(r), s(c)), negated
// when s takes it across the diagonal, and only entries in row or column kp read it from a line.
// Per step: candidates -> pivot -> publish lines -> barrier -> pivot value, tau and column k+1
// at the tile's ranges from the lines, rank-2 update of the own tiles, candidates of column
// k+2 -> barrier.  The arithmetic is that of pfaff_one.
// ------------------------------------------------------------------------------------------
constexpr int CTB = 64;

struct ZCoopArgs {
    const double2* A;   // n x n, ld n (read below the diagonal)
    double2* L;         // n x n work copy, ld n
    double2* W;         // 3 lines of n
    double* part;       // 2 G words: candidate modulus, row
    unsigned* bar;
    int* err;
    double* o1;         // mode 0: Pf (2 doubles); mode 1: log|Pf|
    double* o2;         // mode 1: cis(phase) (2 doubles)
    int n;
    int mode;
};

// tile k of the list ordered by decreasing block column J (rows I = J .. nt-1 within it)
__device__ __forceinline__ void ltile_of(int64_t k, int nt, int& I, int& J)
{
    int m = (int)((sqrt(8.0 * (double)k + 1.0) - 1.0) * 0.5);
    while ((int64_t)(m + 1) * (m + 2) / 2 <= k) m++;
    while ((int64_t)m * (m + 1) / 2 > k) m--;
    J = nt - 1 - m;
    I = J + (int)(k - (int64_t)m * (m + 1) / 2);
}

__device__ __forceinline__ cplx ldc2(const double2* p)
{
    const double* q = (const double*)p;
    return cplx{ld_co(q), ld_co(q + 1)};
}
__device__ __forceinline__ void stc2(double2* p, cplx v)
{
    double* q = (double*)p;
    st_co(q, v.re);
    st_co(q + 1, v.im);
}

struct ZCoopShared {
    cplx taur[CTB], colr[CTB], tauc[CTB], colc[CTB];
    double bv[4];
    int bi[4];
    int kp;
    int ok;
};

// workgroup argmax (Julia's order) into s.kp; every thread calls
__device__ void zcoop_argmax(ZCoopShared& s, double bv, int bi)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) amax_merge(bv, bi, __shfl_xor(bv, o, 64), __shfl_xor(bi, o, 64));
    if (lane == 0) {
        s.bv[wave] = bv;
        s.bi[wave] = bi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double v = s.bv[0];
        int i = s.bi[0];
        for (int w = 1; w < 4; w++) amax_merge(v, i, s.bv[w], s.bi[w]);
        s.bv[0] = v;
        s.kp = i;
    }
    __syncthreads();
}

__global__ __launch_bounds__(PW) void k_zpfaff_coop(ZCoopArgs g)
{
    __shared__ ZCoopShared s;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int n = g.n, G = gridDim.x, gi = blockIdx.x;
    const int nt = (n + CTB - 1) / CTB;
    const int64_t T = (int64_t)nt * (nt + 1) / 2;
    double2* __restrict__ L = g.L;
    unsigned epoch = 0;

    // ---- copy the own tiles (strictly lower part), candidates of column 0 ----
    double bv = -1.0;
    int bi = n;
    for (int64_t k = gi; k < T; k += G) {
        int I, J;
        ltile_of(k, nt, I, J);
        const int r = I * CTB + lane;
        for (int cc = wave; cc < CTB; cc += PW / 64) {
            const int c = J * CTB + cc;
            if (r < n && c < n && r > c) {
                const double2 v = g.A[r + (int64_t)c * n];
                L[r + (int64_t)c * n] = v;
                if (c == 0) amax_merge(bv, bi, hypot(v.x, v.y), r);
            }
        }
    }
    zcoop_argmax(s, bv, bi);
    if (tid == 0) {
        st_co(g.part + 2 * gi, s.bv[0]);
        st_co(g.part + 2 * gi + 1, (double)s.kp);
    }
    if (!grid_sync_co(g.bar, ++epoch, g.err, &s.ok)) return;

    cplx pf = cplx{1.0, 0.0};
    double logpf = 0.0, phase = 0.0;
    bool zero = false;
    for (int k = 0; k + 1 < n; k += 2) {
        // ---- pivot row from every workgroup's candidate of column k ----
        {
            double v = -1.0;
            int i = n;
            for (int q = tid; q < G; q += PW) amax_merge(v, i, ld_co(g.part + 2 * q), (int)ld_co(g.part + 2 * q + 1));
            zcoop_argmax(s, v, i);
        }
        const int kp = s.kp;
        const int k1 = k + 1;
        const int P[3] = {k, k1, kp};
        auto slot = [&](int m) { return m == k ? 0 : m == k1 ? 1 : 2; };
        auto sig = [&](int a) { return a == k1 ? kp : a == kp ? k1 : a; };
        // ---- publish the lines of k, k+1 and kp (entries at t >= k) ----
        for (int64_t q = gi; q < T; q += G) {
            int I, J;
            ltile_of(q, nt, I, J);
            for (int e = 0; e < 3; e++) {
                const int m = P[e];
                if (slot(m) != e) continue;
                double2* Wm = g.W + (int64_t)e * n;
                if (m / CTB == J)   // column m: rows t > m
                    for (int t = I * CTB + tid; t < min(I * CTB + CTB, n); t += PW)
                        if (t > m) stc2(Wm + t, ld2(L + t + (int64_t)m * n));
                if (m / CTB == I)   // row m: columns k <= t < m
                    for (int t = J * CTB + tid; t < min(J * CTB + CTB, n); t += PW)
                        if (t < m && t >= k) stc2(Wm + t, ld2(L + m + (int64_t)t * n));
            }
        }
        if (!grid_sync_co(g.bar, ++epoch, g.err, &s.ok)) return;
        // new entry S[x, y] (x != y, x, y >= k) with s(x) or s(y) in {k, k+1, kp}
        auto snew = [&](int x, int y) {
            const int u = sig(x), v = sig(y);
            const bool lo = u > v;
            const int hi = lo ? u : v, low = lo ? v : u;
            const cplx w = (hi == k || hi == k1 || hi == kp) ? ldc2(g.W + (int64_t)slot(hi) * n + low)
                                                             : ldc2(g.W + (int64_t)slot(low) * n + hi);
            return lo ? w : cneg(w);
        };
        const cplx sub = snew(k1, k);   // A[k+1, k] after the swap
        if (kp != k1) {
            if (g.mode == 0) pf = cneg(pf);
            else phase += 3.14159265358979323846;   // Float64(pi)
        }
        if (sub.re == 0.0 && sub.im == 0.0) {
            zero = true;
            break;   // the same value in every workgroup
        }
        const cplx a = cneg(sub);
        if (g.mode == 0) pf = cmul(pf, a);
        else {
            logpf += log(hypot(a.re, a.im));
            phase += atan2(a.im, a.re);
        }
        // ---- own tiles: rank-2 update of the trailing part, candidates of column k+2 ----
        const int t0 = k + 2;
        bv = -1.0;
        bi = n;
        const int J0 = k / CTB;
        const int64_t K = (int64_t)(nt - J0) * (nt - J0 + 1) / 2;
        for (int64_t q = gi; q < K; q += G) {
            int I, J;
            ltile_of(q, nt, I, J);
            const int r0 = I * CTB, c0 = J * CTB;
            if (r0 + CTB <= t0 + 1 || c0 + CTB <= t0) continue;   // no entry r > c >= k+2 here (uniform)
            for (int t = tid; t < 2 * CTB; t += PW) {
                const int x = (t < CTB ? r0 : c0) + (t & (CTB - 1));
                cplx tv = cplx{0.0, 0.0}, cv = cplx{0.0, 0.0};
                if (x >= t0 && x < n) {
                    tv = cdiv(cneg(snew(x, k)), a);
                    cv = snew(x, k1);
                }
                if (t < CTB) {
                    s.taur[t] = tv;
                    s.colr[t] = cv;
                } else {
                    s.tauc[t - CTB] = tv;
                    s.colc[t - CTB] = cv;
                }
            }
            __syncthreads();
            const int r = r0 + lane;
            for (int cc = wave; cc < CTB; cc += PW / 64) {
                const int c = c0 + cc;
                if (r >= n || c < t0 || r <= c) continue;
                double2* e = L + r + (int64_t)c * n;
                const cplx old = (r == kp || c == kp) && kp != k1 ? snew(r, c) : ld2(e);
                const cplx d = csub(cmul(s.taur[lane], s.colc[cc]), cmul(s.tauc[cc], s.colr[lane]));
                const cplx w = cadd(old, d);
                st2(e, w);
                if (c == t0) amax_merge(bv, bi, hypot(w.re, w.im), r);
            }
            __syncthreads();
        }
        zcoop_argmax(s, bv, bi);
        if (tid == 0) {
            st_co(g.part + 2 * gi, s.bv[0]);
            st_co(g.part + 2 * gi + 1, (double)s.kp);
        }
        if (!grid_sync_co(g.bar, ++epoch, g.err, &s.ok)) return;
    }
    if (gi == 0 && tid == 0) {
        if (g.mode == 0) {
            g.o1[0] = zero ? 0.0 : pf.re;
            g.o1[1] = zero ? 0.0 : pf.im;
        } else {
            g.o1[0] = zero ? -INFINITY : logpf;
            g.o2[0] = zero ? 0.0 : cos(phase);
            g.o2[1] = zero ? 0.0 : sin(phase);
        }
    }
}

}  // namespace zpf
}  // namespace skh

using namespace skh;

extern "C" int skh_internal_zpfaff_coop(const double2* A, double2* L, int64_t n, int mode, double* o1, double* o2,
                                        double2* W, double* part, unsigned* bar, int* err, int grid, hipStream_t st)
{
    zpf::ZCoopArgs a;
    a.A = A;
    a.L = L;
    a.W = W;
    a.part = part;
    a.bar = bar;
    a.err = err;
    a.o1 = o1;
    a.o2 = o2;
    a.n = (int)n;
    a.mode = mode;
    void* args[] = {(void*)&a};
    return (int)hipLaunchCooperativeKernel((const void*)zpf::k_zpfaff_coop, dim3(grid), dim3(zpf::PW), args, 0, st);
}

extern "C" int skh_internal_zpfaff_coop_occupancy(int* per_cu)
{
    return (int)hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, (const void*)zpf::k_zpfaff_coop, zpf::PW, 0);
}

extern "C" int skh_internal_zpfaff(const double2* A, const int64_t* ns, const int64_t* offs, const int64_t* order,
                                   int64_t batch, int64_t nmax, double2* slots, int64_t slot, int grid, int mode, double* o1,
                                   double* o2, hipStream_t st)
{
    if (batch <= 0) return 0;
    if (nmax <= zpf::NSMALL)
        hipLaunchKernelGGL(zpf::k_zpfaff<zpf::NSMALL>, dim3((unsigned)grid), dim3(zpf::PW), 0, st, A, ns, offs, order,
                           batch, slots, slot, mode, o1, o2);
    else
        hipLaunchKernelGGL(zpf::k_zpfaff<zpf::NMAX>, dim3((unsigned)grid), dim3(zpf::PW), 0, st, A, ns, offs, order,
                           batch, slots, slot, mode, o1, o2);
    return (int)hipGetLastError();
}

// workgroups per CU of the variant that serves matrices up to nmax
extern "C" int skh_internal_zpfaff_occupancy(int64_t nmax, int* per_cu)
{
    const void* f = nmax <= zpf::NSMALL ? (const void*)zpf::k_zpfaff<zpf::NSMALL> : (const void*)zpf::k_zpfaff<zpf::NMAX>;
    return (int)hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, f, zpf::PW, 0);
}

extern "C" int skh_internal_zpfaff_nmax() { return zpf::NMAX; }
