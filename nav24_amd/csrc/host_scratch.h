// host_scratch.h — one device buffer per device for the host-array entry
// points (the *_host functions): grown to the largest call and kept until the
// process ends, so a call does not pay hipMalloc / hipFree.  A call holds the
// device's mutex from its first copy in to its last copy out; those calls are
// synchronous, so the buffer is free again when the mutex is released.
#ifndef NAV24_HOST_SCRATCH_H
#define NAV24_HOST_SCRATCH_H

#include <hip/hip_runtime.h>

#include <cstddef>
#include <mutex>
#include <vector>

namespace navba {

struct HostScratch {
    std::mutex mu;
    void* p = nullptr;
    size_t cap = 0;
};

inline HostScratch& host_scratch(int device) {
    static std::mutex g;
    static std::vector<HostScratch*> v;
    std::lock_guard<std::mutex> lk(g);
    if ((int)v.size() <= device) v.resize(device + 1, nullptr);
    if (!v[device]) v[device] = new HostScratch();
    return *v[device];
}

// At least `bytes` on the current device; the caller holds s.mu.  nullptr if
// the allocation fails.
inline void* host_scratch_get(HostScratch& s, size_t bytes) {
    if (s.cap < bytes) {
        if (s.p) (void)hipFree(s.p);
        s.p = nullptr;
        s.cap = 0;
        if (hipMalloc(&s.p, bytes) != hipSuccess) {
            s.p = nullptr;
            return nullptr;
        }
        s.cap = bytes;
    }
    return s.p;
}

}  // namespace navba

#endif  // NAV24_HOST_SCRATCH_H
