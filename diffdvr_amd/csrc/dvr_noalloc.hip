// Allocation-free variant of dvr_render_fused: the checkpoint scratch of the
// fused forward + loss + checkpointed adjoint lives in a file-scope device
// array of DVR_NOALLOC_SCRATCH_FLOATS floats instead of the library's
// grow-only workspace, so a call allocates and frees nothing (for callers
// that cannot: HIP graph capture, tight serving loops).
//
// Every GPU the library runs on holds its own copy of the array (32 MiB) until
// the process ends.  Host code never passes the array's name to a kernel (the
// host-side symbol is a stand-in, not a device address): the device address
// comes from hipGetSymbolAddress, once per device with that device current,
// at the entry point's first call on that device (the runtime then loads the
// array: a caller that captures graphs makes that first call before capturing),
// and is kept per device.  Calls that overlap on one GPU (two streams) share
// the array and race; the checkpoints are written before they are read within
// one call, so nothing needs clearing between calls.
#include <hip/hip_runtime.h>

#include <mutex>
#include <vector>

#include "../../include/dvr.h"
#include "dvr_host_util.hpp"

static __device__ __align__(256) float scratch[DVR_NOALLOC_SCRATCH_FLOATS];

namespace {

std::mutex g_scratchMu;
std::vector<float*> g_scratchPtr;  // device address of `scratch` per device ordinal (NULL = not yet)

// Called by the fused entry point with `device` current.
float* scratch_of(int device)
{
    std::lock_guard<std::mutex> lk(g_scratchMu);
    if (device < 0) dvr_host::fail(DVR_ERR_UNSUPPORTED, "the allocation-free entry point renders on the GPU only");
    if ((int)g_scratchPtr.size() <= device) g_scratchPtr.resize(device + 1, nullptr);
    if (!g_scratchPtr[device]) {
        int cur = -1;
        dvr_host::check_hip(hipGetDevice(&cur), "hipGetDevice");
        if (cur != device)
            dvr_host::fail(DVR_ERR_HIP, "scratch set-up: device %d is not current (device %d is)", device, cur);
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            dvr_host::fail(DVR_ERR_HIP, "hipGetSymbolAddress(scratch) failed on device %d: %s", device,
                           hipGetErrorString(e));
        g_scratchPtr[device] = static_cast<float*>(p);
    }
    return g_scratchPtr[device];
}

}  // namespace

extern "C" dvr_status dvr_render_fused_noalloc(const dvr_inputs* inputs, dvr_outputs* outputs, const dvr_loss* loss,
                                               dvr_adjoint_outputs* adj_outputs, int32_t checkpoint_interval,
                                               void* stream)
{
    const dvr_host::FixedScratch fixed{&scratch_of, (int64_t)DVR_NOALLOC_SCRATCH_FLOATS};
    return dvr_host::fused_fixed_scratch(inputs, outputs, loss, adj_outputs, checkpoint_interval, stream, fixed);
}
