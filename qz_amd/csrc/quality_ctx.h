// quality_ctx coder (host): qz's context-adaptive quality compressor,
// crates/qz-lib/src/compression/quality_ctx.rs.
#pragma once
#include <cstddef>
#include <cstdint>
#include <vector>

namespace qz {
namespace qctx {

// Reads are concatenated: read r's qualities are qlen[r] bytes (Phred+33),
// its sequence slen[r] bytes. Returns 0 (blob filled; empty for no reads)
// or QZ_BAD_PARAMETER (quality byte < 33, no symbols, or a quality string
// longer than its sequence).
int compress(const uint8_t* quals, const int* qlen, const uint8_t* seqs, const int* slen, int nreads,
             std::vector<uint8_t>& blob);

// Decodes nreads reads; quals must hold the sum of their lengths (the
// stored read length, or each sequence's length for variable-length blobs),
// qlen receives them. Returns 0 or QZ_DATA_CORRUPT.
int decompress(const uint8_t* data, size_t size, const uint8_t* seqs, const int* slen, int nreads, uint8_t* quals,
               int* qlen);

}  // namespace qctx
}  // namespace qz
