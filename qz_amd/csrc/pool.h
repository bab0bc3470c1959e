// Pool of device engines. Each slot owns one stream (default priority) and
// the resident workspaces of both BWT directions; a caller holds a slot for
// the duration of one call, so calls from different host threads overlap on
// the device without sharing workspace.
#pragma once
#include "engine.h"
#include "qlfc_gpu.h"

namespace qz {

struct Slot {
    BwtEngine fwd;
    IbwtEngine inv;
    QlfcFront front;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool busy = false;
};

// Returns a free slot (creating one if fewer than the pool limit exist),
// blocking while all are busy; nullptr if the device cannot be initialised.
Slot* pool_acquire();
void pool_release(Slot* s);
int pool_set_limit(int n);
// Releases the workspaces of every idle slot and the cached pinned host
// buffers; returns their footprint in bytes.
uint64_t pool_trim();

// Pinned host buffers for device-to-host downloads, cached across calls
// (pinning is slow, and DMA into pinned memory is several times faster than
// into pageable memory). pinned_acquire returns at least `bytes` bytes and
// their capacity in *cap, or nullptr if no memory can be pinned;
// pinned_release returns a buffer to the cache (up to 4 GiB cached).
uint8_t* pinned_acquire(size_t bytes, size_t* cap);
void pinned_release(uint8_t* p, size_t cap);
int pool_device_ok();

class SlotGuard {
public:
    SlotGuard() : s_(pool_acquire()) {}
    ~SlotGuard() { if (s_) pool_release(s_); }
    bool ok() const { return s_ != nullptr; }
    Slot& slot() { return *s_; }
    SlotGuard(const SlotGuard&) = delete;
    SlotGuard& operator=(const SlotGuard&) = delete;

private:
    Slot* s_;
};

}  // namespace qz
