// Serving loop with continuous batching for the BWT stages.
//
// Jobs (a forward BWT of 1..N blocks for a compress request, an inverse BWT
// for a decompress request) arrive from any host thread. Worker threads each
// own one engine slot (stream + resident SA / rank / LF workspaces); a free
// worker takes the oldest pending job and every further pending job of the
// same direction that fits its batch cap, and runs them as ONE device batch
// (blocks laid end to end). A batch is admitted only while the device
// footprint of all batches in flight stays within the HBM budget, so jobs
// that arrive while others run join the next batch as soon as a worker and
// memory free up. Every job records submit / start / finish times.
//
// Small-job lane: one more worker takes only jobs of at most lane_max bytes
// (oldest such job first, batched with later ones up to lane_max), so a
// small request does not queue behind batches of large ones; its kernels run
// on the lane's own stream beside theirs.
//
// No counterpart in the reference: qz issues one libbsc call per block from
// rayon workers (crates/qz-lib/src/compression/mod.rs:402-404,
// decompress_impl.rs:117-120); this replaces that fan-out for a service.
#include "api_internal.h"
#include "../../include/qzbsc.h"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

namespace qz {

namespace {

double now_ms()
{
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

struct Job {
    int id = 0;
    int kind = 0;  // 0 forward, 1 inverse
    const uint8_t* in = nullptr;
    uint8_t* out = nullptr;
    std::vector<int> sizes;
    std::vector<int> prim_in;
    int* prim_out = nullptr;
    int* aux_out = nullptr;
    uint64_t bytes = 0;
    double t_submit = 0, t_start = 0, t_done = 0;
    int status = 1;  // 1 pending, 0 ok, < 0 error
    int batch_jobs = 0;
};

uint64_t job_footprint(int kind, uint64_t bytes, uint32_t blocks)
{
    return kind == 0 ? BwtEngine::footprint(bytes, blocks) : IbwtEngine::footprint(bytes, blocks);
}

}  // namespace

class Server {
public:
    // Starts the workers and returns once every one holds its resident
    // forward and inverse workspaces for a full batch (init_rc() != 0 if
    // any could not).
    Server(uint64_t budget, uint64_t max_batch, int workers, uint64_t lane_max)
        : budget_(budget), max_batch_(max_batch)
    {
        trace_ = getenv("QZBSC_TRACE_SERVE") != nullptr;
        for (int w = 0; w < workers; ++w) threads_.emplace_back([this] { work(0); });
        if (lane_max) threads_.emplace_back([this, lane_max] { work(lane_max); });
        const int all = (int)threads_.size();
        std::unique_lock<std::mutex> lk(mu_);
        cv_done_.wait(lk, [&] { return ready_ == all; });
    }
    int init_rc() const { return init_rc_; }
    static uint32_t batch_blocks(uint64_t bytes) { return (uint32_t)(bytes / 65536 + 64); }
    ~Server()
    {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_work_.notify_all();
        for (auto& t : threads_) t.join();
    }

    int submit(std::unique_ptr<Job> j)
    {
        std::lock_guard<std::mutex> lk(mu_);
        const int id = ++next_id_;
        j->id = id;
        j->t_submit = now_ms();
        pending_.push_back(j.get());
        jobs_[id] = std::move(j);
        cv_work_.notify_all();  // the small-job lane may not take it: wake every idle worker
        return id;
    }

    int wait(int id, double* lat_ms, double* queue_ms, int* batch_jobs)
    {
        std::unique_lock<std::mutex> lk(mu_);
        auto it = jobs_.find(id);
        if (it == jobs_.end()) return QZ_BAD_PARAMETER;
        Job* j = it->second.get();
        cv_done_.wait(lk, [&] { return j->status != 1; });
        if (lat_ms) *lat_ms = j->t_done - j->t_submit;
        if (queue_ms) *queue_ms = j->t_start - j->t_submit;
        if (batch_jobs) *batch_jobs = j->batch_jobs;
        const int st = j->status;
        jobs_.erase(it);
        return st;
    }

    uint64_t peak_inflight() const { return peak_; }

private:
    // lane_max == 0: a regular worker (any job, oldest first); otherwise the
    // small-job lane (jobs of at most lane_max bytes).
    void work(uint64_t lane_max)
    {
        const uint64_t cap = lane_max ? lane_max : max_batch_;
        SlotGuard g;
        {
            int rc = g.ok() ? 0 : QZ_GPU_ERROR;
            if (rc == 0) rc = g.slot().fwd.reserve(cap, batch_blocks(cap));
            if (rc == 0) rc = g.slot().inv.reserve(cap, batch_blocks(cap));
            std::lock_guard<std::mutex> lk(mu_);
            if (rc) init_rc_ = rc;
            ++ready_;
        }
        cv_done_.notify_all();
        for (;;) {
            std::vector<Job*> batch;
            uint64_t fp = 0;
            {
                std::unique_lock<std::mutex> lk(mu_);
                auto eligible = [&](const Job* j) { return !lane_max || j->bytes <= lane_max; };
                std::deque<Job*>::iterator first;
                for (;;) {
                    if (stop_ && pending_.empty()) return;
                    first = std::find_if(pending_.begin(), pending_.end(), eligible);
                    if (first != pending_.end()) {
                        const uint64_t need = job_footprint((*first)->kind, (*first)->bytes,
                                                            (uint32_t)(*first)->sizes.size());
                        // admit when it fits next to what is in flight (a job bigger than
                        // the whole budget runs alone)
                        if (inflight_ == 0 || inflight_ + need <= budget_) break;
                    }
                    if (stop_ && first == pending_.end()) return;  // what is left is not this lane's
                    cv_work_.wait(lk);
                }
                // continuous batching: the oldest job this worker may take, then
                // later jobs of the same direction while cap and budget allow
                const int kind = (*first)->kind;
                uint64_t bytes = 0;
                uint32_t blocks = 0;
                for (auto it = first; it != pending_.end();) {
                    Job* j = *it;
                    if (j->kind != kind || !eligible(j)) {
                        ++it;
                        continue;
                    }
                    const uint64_t nb = bytes + j->bytes;
                    const uint32_t nbl = blocks + (uint32_t)j->sizes.size();
                    const uint64_t nfp = job_footprint(kind, nb, nbl);
                    if (!batch.empty() && (nb > cap || inflight_ + nfp > budget_ || nb >= 0x7fffffffull))
                        break;
                    batch.push_back(j);
                    bytes = nb;
                    blocks = nbl;
                    fp = nfp;
                    it = pending_.erase(it);
                }
                inflight_ += fp;
                if (inflight_ > peak_) peak_ = inflight_;
            }
            const double t0 = now_ms();
            int rc = g.ok() ? run_batch(g.slot(), batch) : QZ_GPU_ERROR;
            if (trace_) {
                uint64_t nb = 0;
                for (Job* j : batch) nb += j->bytes;
                fprintf(stderr, "qzbsc: worker batch kind %d jobs %zu bytes %llu in %.1f ms (queue wait of the first %.1f ms)\n",
                        batch[0]->kind, batch.size(), (unsigned long long)nb, now_ms() - t0, t0 - batch[0]->t_submit);
            }
            {
                std::lock_guard<std::mutex> lk(mu_);
                inflight_ -= fp;
                const double t1 = now_ms();
                for (Job* j : batch) {
                    j->t_start = t0;
                    j->t_done = t1;
                    j->batch_jobs = (int)batch.size();
                    j->status = rc;
                }
            }
            cv_done_.notify_all();
            cv_work_.notify_all();
        }
    }

    static int run_batch(Slot& sl, const std::vector<Job*>& batch)
    {
        std::vector<uint32_t> offs(1, 0);
        std::vector<int> prim;
        for (Job* j : batch)
            for (size_t b = 0; b < j->sizes.size(); ++b) {
                offs.push_back(offs.back() + (uint32_t)j->sizes[b]);
                if (j->kind == 1) prim.push_back(j->prim_in[b]);
            }
        const uint32_t n = offs.back(), nblk = (uint32_t)offs.size() - 1;
        hipStream_t s = sl.stream;
        if (batch[0]->kind == 0) {
            QZ_TRY(sl.fwd.reserve(n, nblk));
            uint64_t o = 0;
            for (Job* j : batch) {
                QZ_HIP(hipMemcpyAsync(sl.fwd.text_ + o, j->in, j->bytes, hipMemcpyHostToDevice, s));
                o += j->bytes;
            }
            std::vector<int> p(nblk), aux((size_t)nblk * 256);
            QZ_TRY(fwd_batch_device(sl, sl.fwd.text_, sl.fwd.out_, offs, p.data(), aux.data(), nullptr));
            o = 0;
            size_t bi = 0;
            for (Job* j : batch) {
                QZ_HIP(hipMemcpyAsync(j->out, sl.fwd.out_ + o, j->bytes, hipMemcpyDeviceToHost, s));
                for (size_t b = 0; b < j->sizes.size(); ++b, ++bi) {
                    if (j->prim_out) j->prim_out[b] = p[bi];
                    if (j->aux_out) memcpy(j->aux_out + b * 256, aux.data() + bi * 256, 256 * sizeof(int));
                }
                o += j->bytes;
            }
            QZ_HIP(hipStreamSynchronize(s));
        } else {
            QZ_TRY(sl.inv.reserve(n, nblk));
            uint64_t o = 0;
            for (Job* j : batch) {
                QZ_HIP(hipMemcpyAsync(sl.inv.U_ + o, j->in, j->bytes, hipMemcpyHostToDevice, s));
                o += j->bytes;
            }
            QZ_TRY(inv_batch_device(sl, sl.inv.U_, sl.inv.T_, offs, prim.data(), nullptr));
            o = 0;
            for (Job* j : batch) {
                QZ_HIP(hipMemcpyAsync(j->out, sl.inv.T_ + o, j->bytes, hipMemcpyDeviceToHost, s));
                o += j->bytes;
            }
            QZ_HIP(hipStreamSynchronize(s));
        }
        return 0;
    }

    std::mutex mu_;
    std::condition_variable cv_work_, cv_done_;
    std::deque<Job*> pending_;
    std::unordered_map<int, std::unique_ptr<Job>> jobs_;
    std::vector<std::thread> threads_;
    uint64_t budget_, max_batch_;
    bool trace_ = false;
    int ready_ = 0, init_rc_ = 0;
    uint64_t inflight_ = 0, peak_ = 0;
    int next_id_ = 0;
    bool stop_ = false;
};

}  // namespace qz

using namespace qz;

extern "C" {

QZBSC_API void* qzbsc_server_create_lanes(long long hbm_budget, long long max_batch_bytes, int workers,
                                          long long small_lane_bytes)
{
    if (hbm_budget <= 0 || max_batch_bytes <= 0 || workers < 1 || workers > 16 || small_lane_bytes < 0)
        return nullptr;
    if (!pool_device_ok()) return nullptr;
    const uint64_t lane = (uint64_t)std::min(small_lane_bytes, max_batch_bytes);
    pool_set_limit(std::max(4, workers + 3));
    // every worker keeps a forward and an inverse workspace for a full batch
    // resident from the start: cap the batch so that all of them together
    // stay inside the budget
    // idle engines of the pool give their workspaces back first; the budget
    // is then also held to what the device has free
    pool_trim();
    uint64_t budget = (uint64_t)hbm_budget;
    size_t dfree = 0, dtotal = 0;
    if (hipMemGetInfo(&dfree, &dtotal) == hipSuccess && dfree) budget = std::min<uint64_t>(budget, dfree / 10 * 9);
    uint64_t cap = (uint64_t)max_batch_bytes;
    auto resident = [&](uint64_t c) {
        const uint32_t nb = Server::batch_blocks(c);
        uint64_t r = (uint64_t)workers * (BwtEngine::footprint(c, nb) + IbwtEngine::footprint(c, nb));
        if (lane) {
            const uint64_t l = std::min<uint64_t>(lane, c);
            const uint32_t lb = Server::batch_blocks(l);
            r += BwtEngine::footprint(l, lb) + IbwtEngine::footprint(l, lb);
        }
        return r;
    };
    while (cap > (1u << 20) && resident(cap) > budget) cap -= cap / 8;
    Server* s = new Server(budget, cap, workers, std::min<uint64_t>(lane, cap));
    if (s->init_rc() != 0) {
        delete s;
        return nullptr;
    }
    return s;
}

QZBSC_API void* qzbsc_server_create(long long hbm_budget, long long max_batch_bytes, int workers)
{
    return qzbsc_server_create_lanes(hbm_budget, max_batch_bytes, workers, 32ll << 20);
}

QZBSC_API void qzbsc_server_destroy(void* srv) { delete (Server*)srv; }

QZBSC_API int qzbsc_server_submit(void* srv, int kind, const unsigned char* in, unsigned char* out, const int* sizes,
                                  int nblocks, const int* primary_in, int* primary_out, int* aux_out)
{
    if (!srv || (kind != 0 && kind != 1) || !in || !out || !sizes || nblocks < 1) return QZ_BAD_PARAMETER;
    if (kind == 1 && !primary_in) return QZ_BAD_PARAMETER;
    auto j = std::make_unique<Job>();
    j->kind = kind;
    j->in = in;
    j->out = out;
    j->prim_out = primary_out;
    j->aux_out = aux_out;
    uint64_t tot = 0;
    for (int b = 0; b < nblocks; ++b) {
        if (sizes[b] < 1) return QZ_BAD_PARAMETER;
        if (kind == 1 && (primary_in[b] < 1 || primary_in[b] > sizes[b])) return QZ_BAD_PARAMETER;
        j->sizes.push_back(sizes[b]);
        if (kind == 1) j->prim_in.push_back(primary_in[b]);
        tot += (uint64_t)sizes[b];
    }
    if (tot >= 0x7fffffffull) return QZ_BAD_PARAMETER;
    j->bytes = tot;
    return ((Server*)srv)->submit(std::move(j));
}

QZBSC_API int qzbsc_server_wait(void* srv, int job, double* latency_ms, double* queue_ms, int* batch_jobs)
{
    if (!srv) return QZ_BAD_PARAMETER;
    return ((Server*)srv)->wait(job, latency_ms, queue_ms, batch_jobs);
}

QZBSC_API long long qzbsc_server_peak_bytes(void* srv) { return srv ? (long long)((Server*)srv)->peak_inflight() : -1; }

QZBSC_API long long qzbsc_footprint(int kind, long long bytes, int blocks)
{
    if (bytes < 0 || blocks < 1) return -1;
    return (long long)job_footprint(kind, (uint64_t)bytes, (uint32_t)blocks);
}

}  // extern "C"
