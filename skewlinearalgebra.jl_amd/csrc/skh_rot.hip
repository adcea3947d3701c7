// skh_rot.hip -- applies the Givens rotations of the skew-tridiagonal implicit QR to the eigenvector
// halves Qodd and Qeven on the GPU (the O(n^3) part of skewtrieigen_divided!, reference
// src/skeweigen.jl:158-193 with the vector updates of :322-457).
//
// The host runs the QR iteration (O(n^2), sequential) and logs every sweep's rotations; a sweep
// acts on Qodd through the column pairs (k - 1, k), (k, k + 1), ... (one chain, consecutive k) and
// likewise on Qeven.  Each row of a matrix is independent, so one thread owns one row and runs the
// whole log in order.  Along a chain the second output of a rotation is the first input of the
// next, so it stays in a register (the carry) and each rotation costs one load (column k, issued a
// block of 32 ahead: those columns are not written before the chain reaches them) and one store.
// Same arithmetic as the host, products and sums rounded separately, so Qodd and Qeven come out
// bit-identical to the host's.
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace skh {
namespace rot {

constexpr int RW = 64;   // one wave per workgroup: rows spread over as many CUs as possible

// the double held by lane u (u a compile-time constant), to every lane
__device__ __forceinline__ double lane_bcast(double x, int u)
{
    const unsigned long long b = __builtin_bit_cast(unsigned long long, x);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, u);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), u);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
constexpr int RB = 32;   // columns loaded ahead

// Kernel parameters are plain restrict pointers (not a struct) so the compiler knows the
// rotation lists are read-only.
__global__ __launch_bounds__(RW) void k_rot_chains(double* __restrict__ Qo, int64_t ldo,
                                                  const int32_t* __restrict__ ko, const double* __restrict__ cso,
                                                  const int64_t* __restrict__ sego, int64_t nsego,
                                                  double* __restrict__ Qe, int64_t lde,
                                                  const int32_t* __restrict__ ke, const double* __restrict__ cse,
                                                  const int64_t* __restrict__ sege, int64_t nsege, int64_t nrows)
{
#pragma clang fp contract(off)
    const bool odd = blockIdx.y == 0;
    const int lane = threadIdx.x;
    const int64_t r = (int64_t)blockIdx.x * RW + lane;
    // every lane stays to the end (the coefficients are broadcast from lanes 0..RB-1); lanes past
    // the last row work on row 0 and store nothing
    const bool live = r < nrows;
    double* __restrict__ q = (odd ? Qo : Qe) + (live ? r : 0);
    const int64_t ld = odd ? ldo : lde;
    const int32_t* __restrict__ K = odd ? ko : ke;
    const double* __restrict__ CS = odd ? cso : cse;
    const int64_t* __restrict__ seg = odd ? sego : sege;
    const int64_t nseg = odd ? nsego : nsege;
    for (int64_t g = 0; g < nseg; g++) {
        const int64_t t0 = seg[g], len = seg[g + 1] - t0;
        const int64_t k0 = K[t0];   // columns k0 - 1, k0, ..., k0 + len - 1
        double carry = q[(k0 - 1) * ld];
        auto step = [&](int64_t t, double c, double s, double om) {
            const double sg = carry;
            const double out = c * sg + s * om;
            if (live) q[(k0 + (t - t0) - 1) * ld] = out;
            carry = -s * sg + c * om;
        };
        int64_t j0 = 0;
        if (len >= RB) {
            // full blocks, each with the next block's columns and coefficients in flight (no branch
            // inside a block, so the waits count outstanding accesses instead of draining them); the
            // coefficients of rotation u of a block sit in lane u and are broadcast with readlane
            double cur[RB], nxt[RB];
            double cc = lane < RB ? CS[2 * (t0 + lane)] : 0.0, sc = lane < RB ? CS[2 * (t0 + lane) + 1] : 0.0;
            double cn = 0.0, sn = 0.0;
#pragma unroll
            for (int u = 0; u < RB; u++) cur[u] = q[(k0 + u) * ld];
            for (; j0 + 2 * RB <= len; j0 += RB) {
                // coefficients first: loads retire in order, and the next block needs them first
                if (lane < RB) {
                    cn = CS[2 * (t0 + j0 + RB + lane)];
                    sn = CS[2 * (t0 + j0 + RB + lane) + 1];
                }
#pragma unroll
                for (int u = 0; u < RB; u++) nxt[u] = q[(k0 + j0 + RB + u) * ld];
#pragma unroll
                for (int u = 0; u < RB; u++) step(t0 + j0 + u, lane_bcast(cc, u), lane_bcast(sc, u), cur[u]);
#pragma unroll
                for (int u = 0; u < RB; u++) cur[u] = nxt[u];
                cc = cn;
                sc = sn;
            }
#pragma unroll
            for (int u = 0; u < RB; u++) step(t0 + j0 + u, lane_bcast(cc, u), lane_bcast(sc, u), cur[u]);
            j0 += RB;
        }
        for (; j0 < len; j0++) step(t0 + j0, CS[2 * (t0 + j0)], CS[2 * (t0 + j0) + 1], q[(k0 + j0) * ld]);
        if (live) q[(k0 + len - 1) * ld] = carry;
    }
}

}  // namespace rot
}  // namespace skh

using namespace skh;

// w = 0: Qodd, 1: Qeven.  seg: nseg + 1 entries per matrix.
extern "C" int skh_internal_rot_chains(double* Qo, int64_t ldo, const int32_t* ko, const double* cso,
                                       const int64_t* sego, int64_t nsego, double* Qe, int64_t lde,
                                       const int32_t* ke, const double* cse, const int64_t* sege, int64_t nsege,
                                       int64_t nrows, hipStream_t st)
{
    if (nrows <= 0) return 0;
    hipLaunchKernelGGL(rot::k_rot_chains, dim3((unsigned)((nrows + rot::RW - 1) / rot::RW), 2), dim3(rot::RW), 0, st,
                       Qo, ldo, ko, cso, sego, nsego, Qe, lde, ke, cse, sege, nsege, nrows);
    return (int)hipGetLastError();
}
