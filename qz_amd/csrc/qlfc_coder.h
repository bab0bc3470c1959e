// Host side of the QLFC stage: segment split, the per-segment run/rank
// front end (host form; the device form is in qlfc_front.hip), and the
// adaptive binary coder of ranks and run lengths.
#pragma once
#include <cstdint>
#include <vector>

namespace qz {
namespace qlfc {

constexpr int kMaxSegments = 8;

// Front-end output of one segment: per run (forward order) its symbol, rank
// and length, plus the initial symbol table.
struct SegRuns {
    const uint8_t* sym = nullptr;
    const uint8_t* rank = nullptr;
    const uint32_t* len = nullptr;
    uint32_t count = 0;
    uint8_t table[256];
    int nsym = 0;
};

// Number of coder segments for a block of n bytes.
int num_segments(int n);
// Segment boundaries (libbsc's rank-change sampling every 32 bytes).
void split_segments(const uint8_t* bwt, int n, int nb, int* start, int* size);

// Host front end of one segment; arrays hold at least n entries.
uint32_t front_end(const uint8_t* in, int n, uint8_t* sym, uint8_t* rank, uint32_t* len, uint8_t* table,
                   int* nsym);

// Adaptive coding of one segment of n bytes. Returns bytes written, or
// -3 (not compressible) when the output would exceed cap.
int encode_segment(const SegRuns& runs, int n, uint8_t* out, int cap);
// Decodes one coded segment into out[0..n). Returns 0 or a negative error.
int decode_segment(const uint8_t* in, int in_size, uint8_t* out, int n);
// Reads only the size word and the symbol table of a coded segment.
int read_segment_header(const uint8_t* in, int in_size, int* n, uint8_t* table, int* nsym);

}  // namespace qlfc
}  // namespace qz
