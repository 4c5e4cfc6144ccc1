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

}  // namespace dvr_host
