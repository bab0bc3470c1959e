// LZP preprocessing of a block before its BWT (host): libbsc's optional
// stage, selected by bsc_compress's lzpHashSize / lzpMinLen
// (crates/qz-lib/src/compression/bsc.rs:194-196, 237-246).
#pragma once
#include <cstdint>

namespace qz {
namespace lzp {

// Encodes in[0..n) into out (capacity cap). Returns the encoded size, or
// QZ_NOT_COMPRESSIBLE when the result would not be smaller than the input,
// or QZ_BAD_PARAMETER.
int encode(const uint8_t* in, int n, uint8_t* out, int cap, int hash_bits, int min_len);

// Decodes in[0..m) into out (exactly n bytes). Returns 0, or
// QZ_DATA_CORRUPT when the stream does not decode to n bytes.
int decode(const uint8_t* in, int m, uint8_t* out, int n, int hash_bits, int min_len);

bool params_ok(int hash_bits, int min_len);

}  // namespace lzp
}  // namespace qz
