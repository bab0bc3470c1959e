// Block container helpers shared by the drop-in ABI and the batch/serving
// paths.
#pragma once
#include <cstddef>
#include <cstdint>

namespace qz {

constexpr int kHeader = 28;

struct BlockHeader {
    int block_size = 0, data_size = 0, mode = 0, index = 0;
    uint32_t adler_data = 0, adler_body = 0;
};

uint32_t adler32(const uint8_t* p, size_t n);
int bwt_sample_rate(int n);
int store_block(const uint8_t* in, uint8_t* out, int n);
int check_params(int n, int lzpHashSize, int lzpMinLen, int blockSorter, int coder, int* mode);
int coder_compress(const uint8_t* bwt, int n, uint8_t* out, int cap, bool parallel);
int coder_decompress(const uint8_t* in, int in_size, uint8_t* out, int n, bool parallel);
int finish_block(const uint8_t* input, const uint8_t* bwt, int n, int mode, int primary, const int* aux,
                 uint8_t* out, bool parallel);
int parse_header(const uint8_t* h, int size, BlockHeader* bh);
int decode_payload(const uint8_t* input, int inputSize, const BlockHeader& bh, uint8_t* bwt, bool parallel);

}  // namespace qz
