// Host-side helpers shared by the C ABI translation units: error reporting,
// device allocation and the current-device guard.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "../../include/nbg.h"

namespace nbg {

inline thread_local std::string g_err;

inline int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

#define NBG_HIP(call)                                                                             \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            return ::nbg::fail(NBG_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));   \
    } while (0)

template <typename T>
int dalloc(T **p, int64_t count) {
    if (count <= 0) count = 1;
    NBG_HIP(hipMalloc((void **)p, (size_t)count * sizeof(T)));
    return NBG_OK;
}

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        (void)hipSetDevice(dev);
    }
    ~DeviceGuard() {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

}  // namespace nbg

// A context binds one GPU.
struct nbg_ctx {
    int device = 0;
    // this device's copy of the allocation-free passive-point scratch
    // (looked up with hipGetSymbolAddress at nbg_ctx_create)
    void *noalloc_scratch = nullptr;
};
