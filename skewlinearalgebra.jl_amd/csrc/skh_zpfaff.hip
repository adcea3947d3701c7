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

}  // namespace zpf
}  // namespace skh

using namespace skh;

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
