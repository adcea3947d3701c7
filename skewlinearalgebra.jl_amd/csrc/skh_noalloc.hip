// skh_noalloc.hip -- the fixed device scratch of skh_sytrd_d_device_noalloc.
//
// One file-scope device array per GPU holds what the large engine (persistent panel kernel +
// MFMA rank-2k update) otherwise takes from the library's cached buffers: the 2 KiB control
// block (error flag, phase clocks) followed by the workspace (W and V of the panel, the copies of
// the norm and W^T x / V^T x partials and of the column, t, the matvec row and column partials).
// It is sized for n <= SKH_NOALLOC_MAX_N at any panel width up to skh::NBMAX on a 256-CU device.
//
// In host code `scratch` names a host-side stand-in, not device memory: its device address is
// looked up once per device (skh_internal_noalloc_setup, from the library's init with that device
// current) and only that address is handed to kernels.
#include <hip/hip_runtime.h>

#include "skh_internal.h"

namespace {

// control block in doubles (2048 bytes), then the workspace at n = 8192, nb = 128, 514 row-partial
// slots: 2 ld nb + 4 (nt + 2 nt NBMAX + ld) + 3 (2 NBMAX + (nseg + nts) ld) + 64 = 15,706,440
// doubles (ld = 8192, nt = 129, nts = 33)
constexpr size_t kCtlDoubles = 256;
constexpr size_t kScratchDoubles = 15710000;

static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
double* g_scratch_addr[kMaxDevices];   // device address of `scratch` on each device, once looked up

}  // namespace

extern "C" int skh_internal_noalloc_setup(int device)
{
    if (device < 0 || device >= kMaxDevices) return (int)hipErrorInvalidDevice;
    if (g_scratch_addr[device]) return 0;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return (int)e;
    g_scratch_addr[device] = (double*)p;
    return 0;
}

extern "C" int skh_internal_noalloc_scratch(int device, unsigned** ctl, double** ws, size_t* ws_doubles)
{
    if (device < 0 || device >= kMaxDevices || !g_scratch_addr[device]) return (int)hipErrorNotInitialized;
    *ctl = (unsigned*)g_scratch_addr[device];
    *ws = g_scratch_addr[device] + kCtlDoubles;
    *ws_doubles = kScratchDoubles - kCtlDoubles;
    return 0;
}
