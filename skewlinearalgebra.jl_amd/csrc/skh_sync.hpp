// skh_sync.hpp -- grid barrier shared by the cooperative kernels (panel and complex engines).
#pragma once
#include <hip/hip_runtime.h>

namespace skh {

// ---------------------------------------------------------------------------------------
// Grid barrier of the cooperative panel kernel, two levels: workgroups arrive on one of 8 group
// counters (blockIdx % 8, i.e. the XCD they are dispatched to), the last arrival of a group
// arrives on the top counter, and every workgroup polls the top counter.  Counters only grow
// within one factorization; `epoch` is the 1-based index of this barrier in the factorization.
// Memory ordering: every wave drains its own stores, then one release fence per workgroup before
// arriving and one acquire fence after the wait (relaxed polling, so the wait itself does not
// touch the caches).  A bounded spin turns a lost workgroup into an error flag, not a hang.
// Layout of `bar` (unsigned words): [0] top counter, [32*(1+k)] group k counter.
// ---------------------------------------------------------------------------------------
constexpr unsigned SPIN_LIMIT = 1u << 26;

__device__ inline bool grid_sync(unsigned* bar, unsigned epoch, int* err, int* s_ok)
{
    // this wave's global stores have reached the L2 before the workgroup releases them; the
    // memory clobber keeps the compiler from moving any store below the wait
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned G = gridDim.x;
        const unsigned grp = blockIdx.x & 7u;
        const unsigned ngroups = G < 8u ? G : 8u;
        const unsigned gsize = (G - grp + 7u) / 8u;
        __threadfence();   // release this workgroup's writes device-wide
        const unsigned old = atomicAdd(&bar[32u * (1u + grp)], 1u);
        if (old + 1u == epoch * gsize) atomicAdd(&bar[0], 1u);
        unsigned spins = 0;
        int ok = 1;
        while (__hip_atomic_load(&bar[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < epoch * ngroups) {
            __builtin_amdgcn_s_sleep(2);
            if (++spins > SPIN_LIMIT) {
                atomicOr(err, 1);
                ok = 0;
                break;
            }
        }
        if (ok && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) ok = 0;
        __threadfence();   // acquire: later loads of this workgroup see every arrival's writes
        *s_ok = ok;
    }
    __syncthreads();
    return *s_ok != 0;
}

// The same two-level barrier without fences, for kernels whose cross-workgroup data moves only
// through agent-scope atomic loads and stores (sc1: coherent across the XCDs' L2 caches).  Each
// wave's stores are complete (vmcnt 0) before its workgroup arrives, and the wait itself is a
// compiler barrier, so no load after it is served before the arrival count is seen.  A release
// and acquire fence per workgroup would write back and invalidate the XCD's L2 each time
// (about 20 us per barrier at 512 workgroups on MI355X, 3.4 us without).
__device__ inline void grid_arrive_co(unsigned* bar, unsigned epoch)
{
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned G = gridDim.x;
        const unsigned grp = blockIdx.x & 7u;
        const unsigned gsize = (G - grp + 7u) / 8u;
        const unsigned old = __hip_atomic_fetch_add(&bar[32u * (1u + grp)], 1u, __ATOMIC_RELAXED,
                                                    __HIP_MEMORY_SCOPE_AGENT);
        if (old + 1u == epoch * gsize) __hip_atomic_fetch_add(&bar[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Second half of the barrier: work placed between grid_arrive_co and grid_wait_co may load only
// data no other workgroup writes during the launch.
__device__ inline bool grid_wait_co(unsigned* bar, unsigned epoch, int* err, int* s_ok)
{
    if (threadIdx.x == 0) {
        const unsigned G = gridDim.x;
        const unsigned ngroups = G < 8u ? G : 8u;
        unsigned spins = 0;
        int ok = 1;
        while (__hip_atomic_load(&bar[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < epoch * ngroups) {
            __builtin_amdgcn_s_sleep(1);
            if (++spins > SPIN_LIMIT) {
                __hip_atomic_fetch_or(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ok = 0;
                break;
            }
        }
        if (ok && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) ok = 0;
        *s_ok = ok;
    }
    __syncthreads();
    return *s_ok != 0;
}

__device__ inline bool grid_sync_co(unsigned* bar, unsigned epoch, int* err, int* s_ok)
{
    grid_arrive_co(bar, epoch);
    return grid_wait_co(bar, epoch, err, s_ok);
}

}  // namespace skh
