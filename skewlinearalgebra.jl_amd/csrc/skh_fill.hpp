// skh_fill.hpp -- filling the panel kernels' hand-over slots with the empty marker from extra
// 64-thread workgroups of another launch (the rank-2k update before the panel).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "skh_device.hpp"
#include "skh_internal.h"

namespace skh {

// entries of fill workgroup f (64 threads, FILL_CHUNK consecutive entries of one row of one segment)
constexpr int FILL_CHUNK = 1024;
__device__ inline void fill_chunk(const FillSegs& fs, int64_t f)
{
    int k = 0;
    while (k + 1 < fs.nseg && f >= fs.first[k + 1]) k++;
    const int64_t local = f - fs.first[k];
    const int64_t row = local / fs.chunks[k];
    const int64_t lo = fs.lo[k] + (local % fs.chunks[k]) * FILL_CHUNK;
    const int64_t hi = min(fs.hi[k], lo + FILL_CHUNK);
    double* s = fs.p[k] + row * fs.ld[k];
    const double e = empty_slot();
    for (int64_t r = lo + threadIdx.x; r < hi; r += 64) s[r] = e;
}

// block plan of nseg segments (copied from fs into f): returns the number of fill workgroups, or -1
// for an invalid request
inline int64_t fill_plan(const FillSegs* fs, int nseg, FillSegs& f)
{
    f = FillSegs{};
    if (nseg < 0 || nseg > FillSegs::MAXSEG || (nseg > 0 && !fs)) return -1;
    int64_t nfill = 0;
    if (nseg > 0) {
        f = *fs;
        for (int k = 0; k < nseg; k++) {
            f.first[k] = nfill;
            f.chunks[k] = std::max<int64_t>(1, (f.hi[k] - f.lo[k] + FILL_CHUNK - 1) / FILL_CHUNK);
            nfill += f.rows[k] * f.chunks[k];
        }
        f.first[nseg] = nfill;
        f.nseg = nseg;
    }
    return nfill;
}

}  // namespace skh
