// Side streams for running independent kernels of one call concurrently
// (fork from the caller's stream, join back before returning).
#pragma once
#include <hip/hip_runtime.h>

namespace sr {

struct Fork {
    hipStream_t side[3] = {nullptr, nullptr, nullptr};
    hipEvent_t ev_fork = nullptr;
    hipEvent_t ev_join[3] = {nullptr, nullptr, nullptr};
    bool ok = false;
};

// Side streams (default priority) and events of the current device, created
// once per device and kept for the life of the process.
inline Fork &fork_for_device() {
    static Fork forks[64];
    int dev = 0;
    (void)hipGetDevice(&dev);
    Fork &f = forks[dev & 63];
    if (!f.ok) {
        bool good = hipEventCreateWithFlags(&f.ev_fork, hipEventDisableTiming) == hipSuccess;
        for (int i = 0; i < 3 && good; ++i) {
            good = hipStreamCreateWithFlags(&f.side[i], hipStreamNonBlocking) == hipSuccess &&
                   hipEventCreateWithFlags(&f.ev_join[i], hipEventDisableTiming) == hipSuccess;
        }
        f.ok = good;
    }
    return f;
}


}  // namespace sr
