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
// Releases the workspaces of every idle slot; returns their footprint in bytes.
uint64_t pool_trim();
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
