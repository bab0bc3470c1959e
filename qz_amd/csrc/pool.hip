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

uint64_t pool_trim()
{
    std::lock_guard<std::mutex> lk(g_mu);
    uint64_t freed = 0;
    for (auto& s : g_slots) {
        if (s->busy) continue;
        freed += s->fwd.cap_ ? BwtEngine::footprint(s->fwd.cap_, s->fwd.cap_blk_) : 0;
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
