// skh_compact.hip -- kernels and launchers of the compact engine (device code in skh_compact.hpp).
#include "skh_cblock.hpp"
#include "skh_compact.hpp"
#include "skh_internal.h"

namespace skh {

// a value every lane of the wave holds alike, made scalar (branches on it are wave-uniform)
__device__ __forceinline__ int64_t uniform_i64(int64_t x)
{
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)(uint64_t)x);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)((uint64_t)x >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

template <int NMAX>
__global__ __launch_bounds__(cmp::CW, 2) void k_compact(CompactArgs a)
{
    using namespace cmp;
    __shared__ CShared<NMAX> s;
    const int tid = threadIdx.x;
    const unsigned long long t_start = wall_clock64();
    double* slot = a.slots ? a.slots + (int64_t)blockIdx.x * a.slot_stride : nullptr;
    // Every path through the loop body ends in a barrier: a single-thread block (`if (tid == 0)`)
    // directly before the back edge lets the compiler thread the other threads' path around the
    // queue fetch straight to the barrier at the loop head, which makes them re-read a stale item
    // index forever while thread 0 waits for them to leave that inner loop.
    for (;;) {
        if (tid == 0) {
            int64_t q = (int64_t)atomicAdd(a.queue, 1u);
            if (__hip_atomic_load(a.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) q = a.count;
            s.idx = q;
        }
        __syncthreads();
        const int64_t q = uniform_i64(s.idx);
        __syncthreads();
        if (q >= a.count) break;
        const int64_t b = uniform_i64(a.items[q]);
        const int64_t n = uniform_i64(a.ns[b]);
        if (n == 0 || (n & 1)) {
            if (tid == 0) {
                a.logpf[b] = (n == 0) ? 0.0 : -INFINITY;
                a.sgn[b] = (n == 0) ? 1.0 : 0.0;
            }
            __syncthreads();
            continue;
        }
        double* M = a.fac ? a.fac + a.offs[b] : slot;
        double* tau;
        double* eta;
        if (a.taueta) {
            tau = a.taueta + a.toffs[b];
            eta = tau + n;
        } else {
            tau = slot + a.slot_stride - 2 * NMAX;
            eta = tau + NMAX;
        }
        // strictly lower triangle of the input into the work matrix (coalesced column copies)
        const double* src = a.A + a.offs[b];
        for (int64_t c = 0; c + 1 < n; c++)
            for (int64_t r = c + 1 + tid; r < n; r += CW) M[r + c * n] = src[r + c * n];
        __syncthreads();
        if (!tridiag_one<NMAX>(M, n, tau, eta, s, a.err, t_start, a.budget)) {
            if (tid == 0) {
                a.logpf[b] = NAN;
                a.sgn[b] = NAN;
            }
            break;
        }
        // the factor and tau/eta stores of this workgroup are visible to its own loads after the barrier
        logpf_reduce(eta, tau, n, a.logpf + b, a.sgn + b);   // ends with a barrier
    }
}

template __global__ void k_compact<512>(CompactArgs);
template __global__ void k_compact<1024>(CompactArgs);

// Blocked engine: the same persistent queue as k_compact, each matrix reduced by tridiag_blocked
// (panels of NB reflectors, an MFMA rank-2k update per panel).  The working copy is the slot (or
// the caller's factor buffer), the W panel a workspace of its own per workgroup.
template <int NMAX, int NB>
__global__ __launch_bounds__(cbk::CW, 2) void k_cblock(CompactArgs a)
{
    using namespace cbk;
    __shared__ BShared<NMAX, NB> s;
    const int tid = threadIdx.x;
    const unsigned long long t_start = wall_clock64();
    double* slot = a.slots ? a.slots + (int64_t)blockIdx.x * a.slot_stride : nullptr;
    double* Wg = a.wbuf + (int64_t)blockIdx.x * a.wstride;
    auto stop = [&]() -> bool {
        if (tid == 0) {
            int st = 0;
            if (a.budget != 0 && wall_clock64() - t_start > a.budget) {
                atomicOr(a.err, 2);
                st = 1;
            }
            if (__hip_atomic_load(a.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) st = 1;
            s.stop = st;
        }
        __syncthreads();
        const bool r = s.stop != 0;
        __syncthreads();
        return r;
    };
    // every path through the loop body ends in a barrier (see k_compact)
    for (;;) {
        if (tid == 0) {
            int64_t q = (int64_t)atomicAdd(a.queue, 1u);
            if (__hip_atomic_load(a.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) q = a.count;
            s.idx = q;
        }
        __syncthreads();
        const int64_t q = uniform_i64(s.idx);
        __syncthreads();
        if (q >= a.count) break;
        const int64_t b = uniform_i64(a.items[q]);
        const int64_t n = uniform_i64(a.ns[b]);
        if (n == 0 || (n & 1)) {
            if (tid == 0) {
                a.logpf[b] = (n == 0) ? 0.0 : -INFINITY;
                a.sgn[b] = (n == 0) ? 1.0 : 0.0;
            }
            __syncthreads();
            continue;
        }
        double* M = a.fac ? a.fac + a.offs[b] : slot;
        double* tau;
        double* eta;
        if (a.taueta) {
            tau = a.taueta + a.toffs[b];
            eta = tau + n;
        } else {
            tau = slot + a.slot_stride - 2 * NMAX;
            eta = tau + NMAX;
        }
        const double* src = a.A + a.offs[b];
        for (int64_t c = 0; c + 1 < n; c++)
            for (int64_t r = c + 1 + tid; r < n; r += CW) M[r + c * n] = src[r + c * n];
        __syncthreads();
        if (!tridiag_blocked<NMAX, NB>(M, n, Wg, tau, eta, s, stop)) {
            if (tid == 0) {
                a.logpf[b] = NAN;
                a.sgn[b] = NAN;
            }
            break;
        }
        logpf_reduce(eta, tau, n, a.logpf + b, a.sgn + b);   // ends with a barrier
    }
}

template __global__ void k_cblock<256, 8>(CompactArgs);
template __global__ void k_cblock<256, 16>(CompactArgs);
template __global__ void k_cblock<512, 8>(CompactArgs);
template __global__ void k_cblock<512, 16>(CompactArgs);

// log|Pf| and sign of one even-n factorization from its eta, tau (one workgroup)
__global__ __launch_bounds__(WG) void k_logpf_from(const double* eta, const double* tau, int64_t n, double* logpf,
                                                   double* sgn)
{
    logpf_reduce(eta, tau, n, logpf, sgn);
}

}  // namespace skh

using namespace skh;

static const void* compact_kernel_for(int nmax) { return nmax <= 512 ? (const void*)k_compact<512> : (const void*)k_compact<1024>; }

extern "C" int skh_internal_compact_occupancy(int nmax, int* blocks_per_cu)
{
    return (int)hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, compact_kernel_for(nmax), cmp::CW, 0);
}

extern "C" int skh_internal_launch_compact(const CompactArgs* a, int nmax, int grid, hipStream_t st)
{
    void* args[] = {(void*)a};
    return (int)hipLaunchKernel(compact_kernel_for(nmax), dim3(grid), dim3(cmp::CW), args, 0, st);
}

static const void* cblock_kernel_for(int nmax, int nb)
{
    if (nmax == 256 && nb == 8) return (const void*)k_cblock<256, 8>;
    if (nmax == 256 && nb == 16) return (const void*)k_cblock<256, 16>;
    if (nmax == 512 && nb == 8) return (const void*)k_cblock<512, 8>;
    if (nmax == 512 && nb == 16) return (const void*)k_cblock<512, 16>;
    return nullptr;
}

extern "C" int skh_internal_cblock_occupancy(int nmax, int nb, int* blocks_per_cu)
{
    const void* k = cblock_kernel_for(nmax, nb);
    if (!k) return (int)hipErrorInvalidValue;
    return (int)hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k, cbk::CW, 0);
}

extern "C" int skh_internal_launch_cblock(const CompactArgs* a, int nmax, int nb, int grid, hipStream_t st)
{
    const void* k = cblock_kernel_for(nmax, nb);
    if (!k) return (int)hipErrorInvalidValue;
    void* args[] = {(void*)a};
    return (int)hipLaunchKernel(k, dim3(grid), dim3(cbk::CW), args, 0, st);
}

extern "C" int skh_internal_logpf_from(const double* eta, const double* tau, int64_t n, double* logpf, double* sgn,
                                       hipStream_t st)
{
    hipLaunchKernelGGL(k_logpf_from, dim3(1), dim3(WG), 0, st, eta, tau, n, logpf, sgn);
    return (int)hipGetLastError();
}
