// Side streams for running independent kernels of one call concurrently
// (fork from the caller's stream, join back before returning).
//
// Each device has one set of side streams and events, created on first use
// and kept for the life of the process, and shared by every call on that
// device.  A ForkScope holds the device's mutex from the fork to the join, so
// calls from several host threads never record or wait on each other's
// events; their side-stream work is queued one call after the other (each
// call's side kernels wait for its own caller stream only).  Once forked, the
// join is always enqueued, also when a launch in between fails, so no kernel
// on a side stream outlives the call's view of the caller's stream.  (A
// caller that captures a graph makes one eager call on the device first, so
// that the streams are not created during the capture.)
#pragma once
#include <hip/hip_runtime.h>

#include <mutex>

namespace sr {

struct Fork {
    hipStream_t side[3] = {nullptr, nullptr, nullptr};
    hipEvent_t ev_fork = nullptr;
    hipEvent_t ev_join[3] = {nullptr, nullptr, nullptr};
    bool ok = false;
    bool tried = false;
    std::mutex mu;  // held by the ForkScope that uses this fork
};

inline Fork *fork_for_device() {
    static std::mutex registry_mu;
    static Fork forks[64];
    int dev = 0;
    (void)hipGetDevice(&dev);
    Fork &f = forks[dev & 63];
    std::lock_guard<std::mutex> lock(registry_mu);
    if (!f.tried) {
        f.tried = true;
        bool good = hipEventCreateWithFlags(&f.ev_fork, hipEventDisableTiming) == hipSuccess;
        for (int i = 0; i < 3 && good; ++i) {
            good = hipStreamCreateWithFlags(&f.side[i], hipStreamNonBlocking) == hipSuccess &&
                   hipEventCreateWithFlags(&f.ev_join[i], hipEventDisableTiming) == hipSuccess;
        }
        f.ok = good;
    }
    return f.ok ? &f : nullptr;
}

// Fork/join of one call.  fork() makes the side streams wait for the
// caller's stream; finish(r) enqueues the join (whether or not r reports an
// error) and returns the first error (r, else a failing join).
class ForkScope {
  public:
    explicit ForkScope(hipStream_t caller) : s_(caller) {}
    ForkScope(const ForkScope &) = delete;
    ForkScope &operator=(const ForkScope &) = delete;
    ~ForkScope() { (void)finish(0); }

    // 0 on success; on failure nothing stays forked.
    int fork() {
        f_ = fork_for_device();
        if (f_ == nullptr) return (int)hipErrorOutOfMemory;
        lock_ = std::unique_lock<std::mutex>(f_->mu);
        int r = (int)hipEventRecord(f_->ev_fork, s_);
        if (r) {
            lock_.unlock();
            f_ = nullptr;
            return r;
        }
        forked_ = true;
        for (int i = 0; i < 3; ++i) {
            const int w = (int)hipStreamWaitEvent(f_->side[i], f_->ev_fork, 0);
            if (w && !r) r = w;
        }
        return r;
    }
    hipStream_t side(int i) const { return f_->side[i]; }

    // The caller's stream waits for the work queued so far on side stream i
    // (its join event is recorded now; finish records it again at the end).
    int handoff(int i) {
        int r = (int)hipEventRecord(f_->ev_join[i], f_->side[i]);
        if (!r) r = (int)hipStreamWaitEvent(s_, f_->ev_join[i], 0);
        return r;
    }

    int finish(int r) {
        if (!forked_) return r;
        forked_ = false;
        for (int i = 0; i < 3; ++i) {
            int j = (int)hipEventRecord(f_->ev_join[i], f_->side[i]);
            if (!j) j = (int)hipStreamWaitEvent(s_, f_->ev_join[i], 0);
            if (j && !r) r = j;
        }
        lock_.unlock();
        return r;
    }

  private:
    hipStream_t s_;
    Fork *f_ = nullptr;
    bool forked_ = false;
    std::unique_lock<std::mutex> lock_;
};

}  // namespace sr
