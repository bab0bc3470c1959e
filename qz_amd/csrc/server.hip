// Serving loop with continuous batching for the BWT stages.
//
// Jobs (a forward BWT of 1..N blocks for a compress request, an inverse BWT
// for a decompress request) arrive from any host thread. Each job is cut at
// submit into work units of whole blocks, at most unit_max bytes each (a
// block larger than that is a unit of its own). Worker threads each own one
// engine slot (stream + resident SA / rank / LF workspaces); a free worker
// takes the first pending unit in scheduling order and every further pending
// unit of the same direction that fits its batch cap, and runs them as ONE
// device batch (blocks laid end to end). Units of one job can run on several
// workers at once. A batch is admitted only while the device footprint of
// all batches in flight stays within the HBM budget, so work that arrives
// while other work runs joins the next batch as soon as a worker and memory
// free up. A job is done when its last unit is.
//
// Scheduling order: oldest job first (FIFO), or the job with the fewest
// bytes not yet started first (shortest remaining, SRPT). Units bound how
// long a newly arrived job waits for a worker to come free.
//
// Small-job lane: one more worker takes only units of jobs of at most
// lane_max bytes, so a small request does not queue behind batches of large
// ones; its kernels run on the lane's own stream beside theirs.
//
// Resident memory: every worker reserves its forward and inverse workspaces
// for a full batch (cap bytes, batch_blocks(cap) blocks) at start. A batch
// never holds more bytes or blocks than that, so the workspaces never grow
// past what server creation allocated; a job with a block larger than the
// cap is refused at submit.
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
#include <list>
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

struct Job;

// A run of whole blocks of one job: blocks [b0, b0 + nb), bytes at offset
// `off` of the job's input and output.
struct Unit {
    Job* job = nullptr;
    int b0 = 0, nb = 0;
    uint64_t off = 0, bytes = 0;
};

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
    uint64_t unstarted = 0;  // bytes of units not yet taken by a worker
    int units_left = 0;  // units not yet finished; the job is done at 0
    double t_submit = 0, t_start = -1, t_done = 0;
    int status = 0;  // 0 ok, < 0 the first error of one of its batches
    int batch_jobs = 0;
    bool waited = false;
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
    Server(uint64_t budget, uint64_t max_batch, int workers, uint64_t lane_max, uint64_t unit_max, int policy)
        : budget_(budget), max_batch_(max_batch), unit_max_(unit_max), policy_(policy)
    {
        // the forward workspace holds one chunk of whole blocks: blocks up
        // to min(batch cap, chunk) are served
        const uint64_t ch = BwtEngine::chunk();
        max_block_ = ch ? std::min(max_batch_, ch) : max_batch_;
        trace_ = getenv("QZBSC_TRACE_SERVE") != nullptr;
        for (int w = 0; w < workers; ++w) threads_.emplace_back([this] { work(0); });
        if (lane_max) threads_.emplace_back([this, lane_max] { work(lane_max); });
        const int all = (int)threads_.size();
        std::unique_lock<std::mutex> lk(mu_);
        cv_done_.wait(lk, [&] { return ready_ == all; });
    }
    int init_rc() const { return init_rc_; }
    static uint32_t batch_blocks(uint64_t bytes) { return (uint32_t)(bytes / 65536 + 64); }
    uint64_t max_batch() const { return max_batch_; }
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
        for (int sz : j->sizes)
            if ((uint64_t)sz > max_block_) return QZ_NOT_ENOUGH_MEMORY;  // a block beyond the workspace chunk
        // cut into units of whole blocks
        std::vector<Unit> units;
        uint64_t off = 0;
        for (int b = 0; b < (int)j->sizes.size();) {
            Unit u;
            u.job = j.get();
            u.b0 = b;
            u.off = off;
            while (b < (int)j->sizes.size() &&
                   (u.nb == 0 || (u.bytes + (uint64_t)j->sizes[b] <= unit_max_ &&
                                  (uint32_t)u.nb < batch_blocks(max_batch_)))) {
                u.bytes += (uint64_t)j->sizes[b];
                ++u.nb;
                ++b;
            }
            if (u.bytes > max_batch_) return QZ_NOT_ENOUGH_MEMORY;  // one block beyond the batch cap
            off += u.bytes;
            units.push_back(u);
        }
        std::lock_guard<std::mutex> lk(mu_);
        const int id = ++next_id_;
        j->id = id;
        j->t_submit = now_ms();
        j->units_left = (int)units.size();
        j->unstarted = j->bytes;
        for (const Unit& u : units) pending_.push_back(u);
        jobs_[id] = std::move(j);
        cv_work_.notify_all();  // the small-job lane may not take it: wake every idle worker
        return id;
    }

    int wait(int id, double* lat_ms, double* queue_ms, int* batch_jobs)
    {
        std::unique_lock<std::mutex> lk(mu_);
        auto it = jobs_.find(id);
        if (it == jobs_.end() || it->second->waited) return QZ_BAD_PARAMETER;  // unknown, or waited on already
        Job* j = it->second.get();
        j->waited = true;
        cv_done_.wait(lk, [&] { return j->units_left == 0; });
        if (lat_ms) *lat_ms = j->t_done - j->t_submit;
        if (queue_ms) *queue_ms = j->t_start - j->t_submit;
        if (batch_jobs) *batch_jobs = j->batch_jobs;
        const int st = j->status;
        jobs_.erase(it);
        return st;
    }

    uint64_t peak_inflight() const { return peak_; }

private:
    // Scheduling order of two units: FIFO by job, or fewest unstarted bytes
    // of their jobs first; units of one job in block order.
    bool before(const Unit& a, const Unit& b) const
    {
        if (a.job != b.job && policy_ == 1 && a.job->unstarted != b.job->unstarted)
            return a.job->unstarted < b.job->unstarted;
        if (a.job != b.job) return a.job->id < b.job->id;
        return a.b0 < b.b0;
    }

    // lane_max == 0: a regular worker (any unit); otherwise the small-job
    // lane (units of jobs of at most lane_max bytes).
    void work(uint64_t lane_max)
    {
        const uint64_t cap = lane_max ? lane_max : max_batch_;
        const uint32_t cap_blk = batch_blocks(cap);
        SlotGuard g;
        {
            int rc = g.ok() ? 0 : QZ_GPU_ERROR;
            if (rc == 0) rc = g.slot().fwd.reserve(cap, cap_blk, std::min(cap, max_block_));
            if (rc == 0) rc = g.slot().fwd.reserve_depths();  // no allocation once serving
            if (rc == 0) rc = g.slot().inv.reserve(cap, cap_blk);
            std::lock_guard<std::mutex> lk(mu_);
            if (rc) init_rc_ = rc;
            ++ready_;
        }
        cv_done_.notify_all();
        for (;;) {
            std::vector<Unit> batch;
            uint64_t fp = 0;
            {
                std::unique_lock<std::mutex> lk(mu_);
                auto eligible = [&](const Unit& u) {
                    return (!lane_max || u.job->bytes <= lane_max) && u.bytes <= cap && (uint32_t)u.nb <= cap_blk;
                };
                std::list<Unit>::iterator first;
                for (;;) {
                    if (stop_ && pending_.empty()) return;
                    first = pending_.end();
                    for (auto it = pending_.begin(); it != pending_.end(); ++it)
                        if (eligible(*it) && (first == pending_.end() || before(*it, *first))) first = it;
                    if (first != pending_.end()) {
                        const uint64_t need = job_footprint(first->job->kind, first->bytes, (uint32_t)first->nb);
                        // admit when it fits next to what is in flight (a unit
                        // bigger than the whole budget runs alone)
                        if (inflight_ == 0 || inflight_ + need <= budget_) break;
                    }
                    if (stop_ && first == pending_.end()) return;  // what is left is not this lane's
                    cv_work_.wait(lk);
                }
                // continuous batching: the first unit in scheduling order, then
                // further units of the same direction, in that order, while the
                // batch cap, its block count and the budget allow
                const int kind = first->job->kind;
                std::vector<std::list<Unit>::iterator> cand;
                for (auto it = pending_.begin(); it != pending_.end(); ++it)
                    if (it != first && it->job->kind == kind && eligible(*it)) cand.push_back(it);
                std::stable_sort(cand.begin(), cand.end(),
                                 [&](std::list<Unit>::iterator a, std::list<Unit>::iterator b) { return before(*a, *b); });
                cand.insert(cand.begin(), first);
                uint64_t bytes = 0;
                uint32_t blocks = 0;
                for (auto it : cand) {
                    const uint64_t nb = bytes + it->bytes;
                    const uint32_t nbl = blocks + (uint32_t)it->nb;
                    const uint64_t nfp = job_footprint(kind, nb, nbl);
                    if (!batch.empty() && (nb > cap || nbl > cap_blk || inflight_ + nfp > budget_)) continue;
                    batch.push_back(*it);
                    bytes = nb;
                    blocks = nbl;
                    fp = nfp;
                    it->job->unstarted -= it->bytes;
                    pending_.erase(it);
                }
                inflight_ += fp;
                if (inflight_ > peak_) peak_ = inflight_;
                const double t0 = now_ms();
                for (Unit& u : batch)
                    if (u.job->t_start < 0) u.job->t_start = t0;
            }
            const double t0 = now_ms();
            double phase[3] = {0, 0, 0};
            int rc = g.ok() ? run_batch(g.slot(), batch, trace_ ? phase : nullptr) : QZ_GPU_ERROR;
            if (trace_) {
                uint64_t nb = 0;
                for (const Unit& u : batch) nb += u.bytes;
                fprintf(stderr, "qzbsc: worker%s batch kind %d units %zu bytes %llu in %.1f ms "
                                "(upload %.1f, device %.1f, download %.1f)\n",
                        lane_max ? " (lane)" : "", batch[0].job->kind, batch.size(), (unsigned long long)nb,
                        now_ms() - t0, phase[0], phase[1], phase[2]);
            }
            {
                std::lock_guard<std::mutex> lk(mu_);
                inflight_ -= fp;
                const double t1 = now_ms();
                int jobs_in_batch = 0;
                for (size_t i = 0; i < batch.size(); ++i)
                    if (i == 0 || batch[i].job != batch[i - 1].job) ++jobs_in_batch;
                for (const Unit& u : batch) {
                    Job* j = u.job;
                    j->batch_jobs = std::max(j->batch_jobs, jobs_in_batch);
                    if (rc && j->status == 0) j->status = rc;  // first error wins; the job still drains
                    if (--j->units_left == 0) j->t_done = t1;
                }
            }
            cv_done_.notify_all();
            cv_work_.notify_all();
        }
    }

    // One device batch: the units' blocks end to end, up and back down.
    // phase (optional): upload, device and download milliseconds.
    static int run_batch(Slot& sl, const std::vector<Unit>& batch, double* phase)
    {
        std::vector<uint32_t> offs(1, 0);
        std::vector<int> prim;
        const int kind = batch[0].job->kind;
        for (const Unit& u : batch)
            for (int b = u.b0; b < u.b0 + u.nb; ++b) {
                offs.push_back(offs.back() + (uint32_t)u.job->sizes[b]);
                if (kind == 1) prim.push_back(u.job->prim_in[b]);
            }
        const uint32_t n = offs.back(), nblk = (uint32_t)offs.size() - 1;
        hipStream_t s = sl.stream;
        double t = now_ms();
        auto lap = [&](int k) {
            if (!phase) return 0;
            QZ_HIP(hipStreamSynchronize(s));
            const double x = now_ms();
            phase[k] = x - t;
            t = x;
            return 0;
        };
        // no growth: n, nblk within the cap, blocks within the chunk
        QZ_TRY(kind == 0 ? sl.fwd.reserve(n, nblk, BwtEngine::max_block(offs)) : sl.inv.reserve(n, nblk));
        uint8_t* d_in = kind == 0 ? sl.fwd.text_ : sl.inv.U_;
        uint8_t* d_out = kind == 0 ? sl.fwd.out_ : sl.inv.T_;
        uint64_t o = 0;
        for (const Unit& u : batch) {
            QZ_HIP(hipMemcpyAsync(d_in + o, u.job->in + u.off, u.bytes, hipMemcpyHostToDevice, s));
            o += u.bytes;
        }
        QZ_TRY(lap(0));
        if (kind == 0) {
            std::vector<int> p(nblk), aux((size_t)nblk * 256);
            QZ_TRY(fwd_batch_device(sl, d_in, d_out, offs, p.data(), aux.data(), nullptr));
            QZ_TRY(lap(1));
            size_t bi = 0;
            for (const Unit& u : batch)
                for (int b = u.b0; b < u.b0 + u.nb; ++b, ++bi) {
                    if (u.job->prim_out) u.job->prim_out[b] = p[bi];
                    if (u.job->aux_out) memcpy(u.job->aux_out + (size_t)b * 256, aux.data() + bi * 256, 256 * sizeof(int));
                }
        } else {
            QZ_TRY(inv_batch_device(sl, d_in, d_out, offs, prim.data(), nullptr));
            QZ_TRY(lap(1));
        }
        o = 0;
        for (const Unit& u : batch) {
            QZ_HIP(hipMemcpyAsync(u.job->out + u.off, d_out + o, u.bytes, hipMemcpyDeviceToHost, s));
            o += u.bytes;
        }
        QZ_HIP(hipStreamSynchronize(s));
        QZ_TRY(lap(2));
        return 0;
    }

    std::mutex mu_;
    std::condition_variable cv_work_, cv_done_;
    std::list<Unit> pending_;
    std::unordered_map<int, std::unique_ptr<Job>> jobs_;
    std::vector<std::thread> threads_;
    uint64_t budget_, max_batch_, unit_max_;
    uint64_t max_block_ = 0;
    int policy_;
    bool trace_ = false;
    int ready_ = 0, init_rc_ = 0;
    uint64_t inflight_ = 0, peak_ = 0;
    int next_id_ = 0;
    bool stop_ = false;
};

}  // namespace qz

using namespace qz;

extern "C" {

QZBSC_API void* qzbsc_server_create_opts(long long hbm_budget, long long max_batch_bytes, int workers,
                                         long long small_lane_bytes, long long unit_bytes, int policy)
{
    if (hbm_budget <= 0 || max_batch_bytes <= 0 || workers < 1 || workers > 16 || small_lane_bytes < 0 ||
        unit_bytes < 0 || (policy != 0 && policy != 1))
        return nullptr;
    if (!pool_device_ok()) return nullptr;
    const uint64_t lane = (uint64_t)std::min(small_lane_bytes, max_batch_bytes);
    pool_set_limit(std::max(4, workers + 3));
    // every worker keeps a forward and an inverse workspace for a full batch
    // resident from the start: cap the batch so that all of them together
    // stay inside the budget; idle engines of the pool give their workspaces
    // back first, and the budget is also held to what the device has free
    pool_trim();
    uint64_t budget = (uint64_t)hbm_budget;
    size_t dfree = 0, dtotal = 0;
    if (hipMemGetInfo(&dfree, &dtotal) == hipSuccess && dfree) budget = std::min<uint64_t>(budget, dfree / 10 * 9);
    uint64_t cap = std::min<uint64_t>((uint64_t)max_batch_bytes, 0x7fffffffull - 1);  // device batches < 2^31 bytes
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
    const uint64_t unit = unit_bytes ? std::min<uint64_t>((uint64_t)unit_bytes, cap) : cap;
    Server* s = new Server(budget, cap, workers, std::min<uint64_t>(lane, cap), unit, policy);
    if (s->init_rc() != 0) {
        delete s;
        return nullptr;
    }
    return s;
}

QZBSC_API void* qzbsc_server_create_lanes(long long hbm_budget, long long max_batch_bytes, int workers,
                                          long long small_lane_bytes)
{
    return qzbsc_server_create_opts(hbm_budget, max_batch_bytes, workers, small_lane_bytes, 64ll << 20, 0);
}

QZBSC_API void* qzbsc_server_create(long long hbm_budget, long long max_batch_bytes, int workers)
{
    return qzbsc_server_create_lanes(hbm_budget, max_batch_bytes, workers, 32ll << 20);
}

QZBSC_API void qzbsc_server_destroy(void* srv) { delete (Server*)srv; }

QZBSC_API long long qzbsc_server_batch_cap(void* srv) { return srv ? (long long)((Server*)srv)->max_batch() : -1; }

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

QZBSC_API void* qzbsc_host_alloc(long long bytes)
{
    if (bytes <= 0) return nullptr;
    void* p = nullptr;
    return hipHostMalloc(&p, (size_t)bytes, hipHostMallocDefault) == hipSuccess ? p : nullptr;
}

QZBSC_API void qzbsc_host_free(void* p)
{
    if (p) (void)hipHostFree(p);
}

QZBSC_API int qzbsc_device_memory(long long* free_bytes, long long* total_bytes)
{
    size_t f = 0, t = 0;
    if (!pool_device_ok() || hipMemGetInfo(&f, &t) != hipSuccess) return QZ_GPU_ERROR;
    if (free_bytes) *free_bytes = (long long)f;
    if (total_bytes) *total_bytes = (long long)t;
    return 0;
}

}  // extern "C"
