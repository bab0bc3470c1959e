// Replay of fixed launch sequences as HIP graphs.
#pragma once

#include <atomic>
#include <cstdlib>
#include <vector>

#include "common.h"

namespace qz {

// Process-wide counts of graph captures and replays (all engines).
inline std::atomic<uint64_t> g_graph_captures{0};
inline std::atomic<uint64_t> g_graph_replays{0};

// One executable graph per launch sequence. The sequence is captured the
// second time in a row that it is asked for with the same shape key, so a
// one-off shape runs eagerly and never pays for a capture, and replayed while
// the key holds. The key must name everything the launches bake in: sizes,
// device pointers, launch shapes. `prepare` (descriptor uploads the launches
// read) runs eagerly, outside the graph, whenever the graph is not replayed;
// a replay relies on those descriptors being untouched since the capture,
// which holds because any other shape first drops the graph. Capture uses
// thread-local mode: other threads of the process keep working meanwhile.
// QZBSC_GRAPH=0 turns replay off.
class GraphCache {
public:
    GraphCache() = default;
    GraphCache(const GraphCache&) = delete;
    GraphCache& operator=(const GraphCache&) = delete;
    ~GraphCache() { drop(); }

    void drop()
    {
        if (exec_) (void)hipGraphExecDestroy(exec_);
        exec_ = nullptr;
        key_.clear();
    }
    // Forget everything, including the shape seen last (buffers changed).
    void reset()
    {
        drop();
        last_.clear();
    }
    uint64_t replays() const { return replays_; }
    uint64_t captures() const { return captures_; }

    template <class P, class L>
    int run(const std::vector<uint64_t>& key, hipStream_t s, P&& prepare, L&& launch)
    {
        if (!enabled()) {
            reset();
            QZ_TRY(prepare(s));
            return launch(s);
        }
        if (exec_ && key == key_) {
            QZ_HIP(hipGraphLaunch(exec_, s));
            ++replays_;
            g_graph_replays.fetch_add(1, std::memory_order_relaxed);
            return 0;
        }
        drop();
        QZ_TRY(prepare(s));
        if (key != last_) {
            last_ = key;
            return launch(s);
        }
        hipGraph_t g = nullptr;
        QZ_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        const int rc = launch(s);
        const hipError_t ec = hipStreamEndCapture(s, &g);
        if (rc != 0 || ec != hipSuccess || !g) {
            if (g) (void)hipGraphDestroy(g);
            (void)hipGetLastError();
            return rc ? rc : QZ_GPU_ERROR;
        }
        const hipError_t ei = hipGraphInstantiate(&exec_, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess) {
            exec_ = nullptr;
            return QZ_GPU_ERROR;
        }
        key_ = key;
        ++captures_;
        g_graph_captures.fetch_add(1, std::memory_order_relaxed);
        QZ_HIP(hipGraphLaunch(exec_, s));
        return 0;
    }

    static bool enabled()
    {
        const char* e = getenv("QZBSC_GRAPH");
        return !(e && e[0] == '0');
    }

private:
    hipGraphExec_t exec_ = nullptr;
    std::vector<uint64_t> key_;   // shape of exec_
    std::vector<uint64_t> last_;  // shape of the previous call
    uint64_t replays_ = 0, captures_ = 0;
};

}  // namespace qz
