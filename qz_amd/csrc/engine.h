// Device engines for the forward and inverse BWT. Each engine owns its HBM
// workspace, grown on demand and kept resident between calls.
#pragma once
#include "common.h"
#include "graph.h"
#include <vector>

namespace qz {

class BwtEngine {
public:
    ~BwtEngine() { release(); }
    // Staging for batches of n bytes in nblk blocks, and a sort workspace
    // for one chunk of them (whole blocks, at most 512 MiB or the largest
    // block, maxblk: by default as large as the batch).
    int reserve(uint64_t n, uint32_t nblk, uint64_t maxblk = ~0ull);
    // The chunk the sort workspace is sized for (bytes; 0: the whole batch).
    static uint64_t chunk();
    // Largest block of a batch.
    static uint64_t max_block(const std::vector<uint32_t>& offs)
    {
        uint64_t m = 0;
        for (size_t b = 0; b + 1 < offs.size(); ++b) m = m > offs[b + 1] - offs[b] ? m : offs[b + 1] - offs[b];
        return m;
    }
    // The depth arrays of run-length batches, allocated now rather than on
    // the first such batch (for callers that must not allocate mid-run).
    int reserve_depths() { return ensure_depths(); }
    void release();
    // off_[0..nblk] must already hold the batch offsets; text/out are device
    // buffers of n bytes (text_/out_ serve as staging for host callers).
    int encode_device(const uint8_t* text, uint8_t* out, uint32_t n, uint32_t nblk, hipStream_t s);
    // Forward BWT of the batch at offs (reserve() done): idx receives every
    // block's primary and sampled indexes (nblk x 256, the layout of idx_).
    // Blocks of different alphabet classes run as separate sub-batches,
    // staged class by class in text_ / out_ (whichever text and out are
    // not; when text is text_, its contents are overwritten). Returns with
    // the stream's work enqueued up to the last copy into out.
    int encode_batch(const uint8_t* text, uint8_t* out, const std::vector<uint32_t>& offs, hipStream_t s,
                     std::vector<int32_t>& idx);
    // Device bytes reserve(n, nblk, maxblk) allocates (admission control);
    // maxblk 0: blocks no larger than a chunk.
    static uint64_t footprint(uint64_t n, uint32_t nblk, uint64_t maxblk = 0);

    uint8_t* text_ = nullptr;    // input symbols
    uint8_t* out_ = nullptr;     // BWT output
    uint32_t* off_ = nullptr;    // block offsets (nblk + 1)
    int32_t* idx_ = nullptr;     // per block: [0] primary, [k] = 1 + rank(k*r)
    uint32_t* sa_ = nullptr;
    uint32_t* rank_ = nullptr;
    int rounds_ = 0;
    bool trace_ = false;
    uint32_t k0_ = 0;
    uint32_t cb_ = 0;            // char mode of the last batch (SA entries carry the preceding symbol)
    uint64_t cap_ = 0;       // staging bytes (text_, out_) and blocks
    uint32_t cap_blk_ = 0;
    uint64_t wcap_ = 0;      // sort workspace: suffixes of one chunk

private:
    int sort_suffixes(const uint8_t* text, uint32_t n, uint32_t nblk, hipStream_t s);
    int classify(const uint8_t* text, const std::vector<uint32_t>& offs, hipStream_t s,
                 std::vector<std::vector<uint32_t>>& classes);
    int enqueue_round0(const uint8_t* text, uint32_t n, uint32_t nblk, int bits, uint32_t ntiles,
                       uint32_t hist_words, bool timed, uint32_t fuse, int fshape, hipStream_t s);
    void* tiles0_ = nullptr;       // round-0 radix tiles (device; read by the round-0 graph)
    void* ppd_ = nullptr;          // symbol packing of round 0 (device)
    uint32_t* btab_ = nullptr;     // position bucket -> block (device), see BlockMap
    std::vector<uint32_t> btab_host_;
    GraphCache g0_;                // round 0 as one graph
    uint32_t* bufA_ = nullptr;
    uint32_t* bufB_ = nullptr;
    uint32_t* keys0_ = nullptr;    // radix keys, one side of the ping-pong (the other: group pool / sa)
    uint32_t* vals0_ = nullptr;
    uint32_t* dep_[2] = {nullptr, nullptr};  // depth of each unsorted group at its start slot, per round
                                             // parity; allocated for the first batch with run-length keys
    uint32_t* lhist_ = nullptr;    // histograms of the large-group sort
    void* lgrp_ = nullptr;         // large groups of a round in slot order (device)
    int32_t* ldelta_ = nullptr;    // per large group: its slot start minus its packed offset
    uint32_t* lmm_ = nullptr;      // per large group: smallest and largest key of the round
    uint32_t* lmm_host_ = nullptr; // pinned copy of lmm_
    hipEvent_t lev_ = nullptr;     // large-group keys gathered (lmm_host_ filled)
    int ensure_depths();
    uint32_t* tmp_ = nullptr;
    uint32_t* cls_ = nullptr;      // per block: sampled symbols and runs (k_block_class)
    void* runs_ = nullptr;         // copy runs of a class split: into the staging, then back
    uint32_t* hist_ = nullptr;
    uint32_t* brun_ = nullptr;     // run boundaries per block (round-0 key mode)
    uint32_t* cnt_ = nullptr;      // group-list counters (two sets) + overflow flag
    void* groups_[2] = {nullptr, nullptr};
    void* tiles_ = nullptr;        // block-aligned radix tiles (device)
    std::vector<TileDesc> tiles_host_;
  public:
    std::vector<uint32_t> offs_host_;  // block offsets of the current batch (host copy)
  private:
    void* pinned_ = nullptr;
    hipEvent_t tev_[8] = {};       // tuner timing: round-0 sort, large-group sort, fused finish, k_sort_small
};

// Digit-width tuner state: which = 0 round 0, 1 large groups. Returns the
// number of timed trials; choice is 0 while undecided.
int tune_state(int which, int* choice, double* ns_per_elem /* [3]: 8, 9, 11 bits */);
// Any knob of the launch-shape tuner (0 round-0 digit width, 1 large-group
// digit width, 2 fused-finish tile shape, 3 k_sort_small LDS budget): the
// chosen value (-1 while undecided), the candidates and the mean ns per unit
// of work of each. Returns the number of timed trials.
int tune_knob(int which, int* choice, int* ncand, int* cand /* [4] */, double* ns_per_unit /* [4] */);

class IbwtEngine {
public:
    ~IbwtEngine() { release(); }
    int reserve(uint64_t n, uint32_t nblk);
    // Places the workspace for batches of at most n bytes in nblk blocks in
    // the device region [base, base + bytes) instead of allocating it (no
    // U_/T_ staging: the caller passes device input and output), and runs
    // every batch with eager launches (no graph). Returns
    // LIBBSC_NOT_ENOUGH_MEMORY if the region is too small.
    int bind(uint8_t* base, size_t bytes, uint64_t n, uint32_t nblk);
    static size_t bound_bytes(uint64_t n, uint32_t nblk);
    void release();
    // off_[0..nblk] and prim_[0..nblk) must hold the batch; U/T are device
    // buffers of n bytes (U_/T_ serve as staging for host callers).
    int decode_device(const uint8_t* U, uint8_t* T, uint32_t n, uint32_t nblk, const std::vector<uint32_t>& offs,
                      hipStream_t s);
    static uint64_t footprint(uint64_t n, uint32_t nblk);

    uint8_t* U_ = nullptr;
    uint8_t* T_ = nullptr;
    uint32_t* off_ = nullptr;
    uint32_t* prim_ = nullptr;
    uint32_t split_ = 0;  // splitter spacing; 0 picks one by batch size
    uint64_t cap_ = 0;
    uint32_t cap_blk_ = 0;

private:
    uint32_t* E_ = nullptr;        // per row: LF | splitter flag
    uint32_t* fst_ = nullptr;      // per block: F-run start row of each present symbol
    uint8_t* fsym_ = nullptr;      // per block: the present symbols
    uint32_t* fsig_ = nullptr;     // per block: number of present symbols
    uint32_t* hist_ = nullptr;
    uint32_t* tmp_ = nullptr;
    void* tiles_ = nullptr;
    uint32_t* sbase_ = nullptr;
    uint32_t* nxt0_ = nullptr;
    uint32_t* nxt1_ = nullptr;
    uint32_t* len_ = nullptr;
    uint32_t* d0_ = nullptr;
    uint32_t* d1_ = nullptr;
    uint32_t* flag_ = nullptr;
    uint8_t* buf_ = nullptr;       // first symbols of every segment
    uint32_t* resume_ = nullptr;   // row after those symbols
    uint32_t* tails_ = nullptr;    // splitters whose segment is longer than its buffer
    uint32_t* rbase_ = nullptr;    // per block: its first ruler of the list ranking
    uint32_t* rul_ = nullptr;      // rulers' next / rows / distances (five arrays)
    void* pinned_ = nullptr;
    uint64_t cap_tiles_ = 0;
    uint64_t cap_split_ = 0;
    bool borrowed_ = false;        // workspace placed by bind(): not ours to free
    GraphCache graph_;             // the whole launch sequence as one graph
    std::vector<TileDesc> td_host_;
    std::vector<uint32_t> sb_host_;
    std::vector<uint32_t> rb_host_;
};

}  // namespace qz
