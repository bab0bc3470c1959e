// Engine pool (see pool.h).
#include "pool.h"
#include "../../include/qzbsc.h"

#include <condition_variable>
#include <memory>
#include <mutex>
#include <vector>

namespace qz {

namespace {
std::mutex g_mu;
std::condition_variable g_cv;
std::vector<std::unique_ptr<Slot>> g_slots;
int g_limit = 4;
int g_dev_state = 0;  // 0 unknown, 1 ok, -1 no device
}  // namespace

int pool_device_ok()
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_dev_state == 0) {
        int cnt = 0;
        g_dev_state = (hipGetDeviceCount(&cnt) == hipSuccess && cnt > 0) ? 1 : -1;
    }
    return g_dev_state == 1;
}

Slot* pool_acquire()
{
    if (!pool_device_ok()) return nullptr;
    std::unique_lock<std::mutex> lk(g_mu);
    for (;;) {
        for (auto& s : g_slots)
            if (!s->busy) {
                s->busy = true;
                return s.get();
            }
        if ((int)g_slots.size() < g_limit) {
            auto s = std::make_unique<Slot>();
            if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) return nullptr;
            if (hipEventCreate(&s->ev0) != hipSuccess || hipEventCreate(&s->ev1) != hipSuccess) return nullptr;
            s->busy = true;
            g_slots.push_back(std::move(s));
            return g_slots.back().get();
        }
        g_cv.wait(lk);
    }
}

void pool_release(Slot* s)
{
    {
        std::lock_guard<std::mutex> lk(g_mu);
        s->busy = false;
    }
    g_cv.notify_one();
}

namespace {
std::mutex g_pin_mu;
std::vector<std::pair<uint8_t*, size_t>> g_pin_free;  // cached pinned buffers
size_t g_pin_cached = 0;
constexpr size_t kPinnedCache = 4ull << 30;
constexpr size_t kPinnedLarge = 64ull << 20;

// Capacities are powers of two from 1 MiB up to 64 MiB, then multiples of
// 64 MiB, so that buffers are reused across calls of similar sizes.
size_t pinned_round(size_t bytes)
{
    if (bytes > kPinnedLarge) return (bytes + kPinnedLarge - 1) / kPinnedLarge * kPinnedLarge;
    size_t c = 1ull << 20;
    while (c < bytes) c <<= 1;
    return c;
}
}  // namespace

uint8_t* pinned_acquire(size_t bytes, size_t* cap)
{
    {
        std::lock_guard<std::mutex> lk(g_pin_mu);
        int best = -1;
        for (int i = 0; i < (int)g_pin_free.size(); ++i)
            if (g_pin_free[i].second >= bytes && (best < 0 || g_pin_free[i].second < g_pin_free[best].second)) best = i;
        if (best >= 0) {
            auto b = g_pin_free[best];
            g_pin_free.erase(g_pin_free.begin() + best);
            g_pin_cached -= b.second;
            *cap = b.second;
            return b.first;
        }
    }
    const size_t c = pinned_round(bytes);
    void* p = nullptr;
    if (hipHostMalloc(&p, c, hipHostMallocDefault) != hipSuccess || !p) {
        (void)hipGetLastError();
        return nullptr;
    }
    *cap = c;
    return static_cast<uint8_t*>(p);
}

void pinned_release(uint8_t* p, size_t cap)
{
    if (!p) return;
    {
        std::lock_guard<std::mutex> lk(g_pin_mu);
        if (g_pin_cached + cap <= kPinnedCache) {
            g_pin_free.push_back({p, cap});
            g_pin_cached += cap;
            return;
        }
    }
    (void)hipHostFree(p);
}

uint64_t pool_trim()
{
    uint64_t freed = 0;
    {
        std::lock_guard<std::mutex> lk(g_pin_mu);
        for (auto& b : g_pin_free) {
            (void)hipHostFree(b.first);
            freed += b.second;
        }
        g_pin_free.clear();
        g_pin_cached = 0;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    for (auto& s : g_slots) {
        if (s->busy) continue;
        freed += s->fwd.cap_ ? BwtEngine::footprint(s->fwd.cap_, s->fwd.cap_blk_, s->fwd.wcap_) : 0;
        freed += s->inv.cap_ ? IbwtEngine::footprint(s->inv.cap_, s->inv.cap_blk_) : 0;
        s->fwd.release();
        s->inv.release();
        s->front.release();
    }
    return freed;
}

int pool_set_limit(int n)
{
    if (n < 1 || n > 64) return QZ_BAD_PARAMETER;
    std::lock_guard<std::mutex> lk(g_mu);
    g_limit = n;
    return 0;
}

}  // namespace qz

extern "C" QZBSC_API int qzbsc_set_max_engines(int n) { return qz::pool_set_limit(n); }
extern "C" QZBSC_API int qzbsc_gpu_available(void) { return qz::pool_device_ok(); }
extern "C" QZBSC_API long long qzbsc_trim(void) { return (long long)qz::pool_trim(); }
