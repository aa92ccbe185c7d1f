// Host-side device properties used to size launches (cached per process).
#pragma once
#include <hip/hip_runtime.h>

namespace sr {

// Compute units of the current device (256 on MI355X).
inline int num_cus() {
    static int v = 0;
    if (v == 0) {
        int dev = 0, c = 0;
        (void)hipGetDevice(&dev);
        if (hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || c <= 0) c = 256;
        v = c;
    }
    return v;
}

// Largest LDS allocation one workgroup of this library requests: the
// device's per-block limit, capped at 64 KiB (kernels that need more keep
// their tapes in global memory instead).  Every launch checks its request
// against it and returns -3 rather than launching a workgroup that no CU
// could host.
inline size_t lds_limit() {
    static size_t v = 0;
    if (v == 0) {
        int dev = 0, c = 0;
        (void)hipGetDevice(&dev);
        if (hipDeviceGetAttribute(&c, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || c <= 0)
            c = 64 * 1024;
        v = (size_t)c < (size_t)(64 * 1024) ? (size_t)c : (size_t)(64 * 1024);
    }
    return v;
}

// The device's own per-workgroup LDS limit (160 KiB on gfx950), for the
// kernels that are sized to use it.
inline size_t lds_limit_device() {
    static size_t v = 0;
    if (v == 0) {
        int dev = 0, c = 0;
        (void)hipGetDevice(&dev);
        if (hipDeviceGetAttribute(&c, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || c <= 0)
            c = 64 * 1024;
        v = (size_t)c;
    }
    return v;
}

}  // namespace sr
