// Host-side error handling shared by the C-ABI translation units: a failed
// check throws Fail{status} with a thread-local message that dvr_last_error()
// returns; guarded() turns it into the entry point's status code.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <exception>
#include <string>

#include "../../include/dvr.h"

namespace dvr_host {

std::string& last_error();

struct Fail {
    dvr_status code;
};

[[noreturn]] inline void fail(dvr_status code, const char* fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    last_error() = buf;
    throw Fail{code};
}

inline void check_hip(hipError_t e, const char* what)
{
    if (e != hipSuccess) fail(DVR_ERR_HIP, "%s failed: %s", what, hipGetErrorString(e));
}

template <class F> dvr_status guarded(F&& f)
{
    try {
        last_error().clear();
        f();
        return DVR_OK;
    } catch (const Fail& e) {
        return e.code;
    } catch (const std::exception& e) {
        last_error() = e.what();
        return DVR_ERR_HIP;
    }
}

inline const char* dtype_name(int d)
{
    switch (d) {
    case DVR_FLOAT32: return "torch.float32";
    case DVR_INT32: return "torch.int32";
    case DVR_FLOAT64: return "torch.float64";
    case DVR_BOOL: return "torch.bool";
    default: return "?";
    }
}

inline bool given(const dvr_tensor& t) { return t.data != nullptr || t.ndim > 0; }

// CHECK_DIM, CHECK_SIZE, CHECK_DTYPE, CHECK_CUDA of renderer/pytorch_utils.h:5-11
inline void check_dim(const dvr_tensor& t, const char* name, int d)
{
    if (!given(t)) fail(DVR_ERR_INVALID_ARGUMENT, "%s is not defined!", name);
    if (t.ndim != d) fail(DVR_ERR_INVALID_ARGUMENT, "%s must be a tensor with %d dimensions, but has %d", name, d, t.ndim);
}
inline void check_size(const dvr_tensor& t, const char* name, int d, int64_t s)
{
    if (t.sizes[d] != s)
        fail(DVR_ERR_INVALID_ARGUMENT, "%s must have %lld entries at dimension %d, but has %lld entries", name,
             (long long)s, d, (long long)t.sizes[d]);
}
inline void check_dtype(const dvr_tensor& t, const char* name, int dt)
{
    if (t.dtype != dt)
        fail(DVR_ERR_INVALID_ARGUMENT, "%s must be of type %s, but is %s", name, dtype_name(dt), dtype_name(t.dtype));
}
inline void check_device(const dvr_tensor& t, const char* name, int device)
{
    if (t.device != device)
        fail(DVR_ERR_INVALID_ARGUMENT, "All tensors must be on the same device, but %s is on device %d (expected %d)",
             name, t.device, device);
}

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev)
    {
        check_hip(hipGetDevice(&prev), "hipGetDevice");
        if (prev != dev) check_hip(hipSetDevice(dev), "hipSetDevice");
    }
    ~DeviceGuard()
    {
        int cur = -1;
        if (hipGetDevice(&cur) == hipSuccess && cur != prev && prev >= 0) (void)hipSetDevice(prev);
    }
};

// Scratch of the allocation-free fused entry point: `get` returns the calling thread's current
// device's static array (it may fail(); it is called with that device current), `floats` its size.
struct FixedScratch {
    float* (*get)(int device);
    int64_t floats;
};
// dvr_render_fused with the checkpoints in that scratch (dvr_capi.cpp)
dvr_status fused_fixed_scratch(const dvr_inputs* inputs, dvr_outputs* outputs, const dvr_loss* loss,
                               dvr_adjoint_outputs* adj_outputs, int32_t checkpoint_interval, void* stream,
                               const FixedScratch& fixed);

}  // namespace dvr_host
