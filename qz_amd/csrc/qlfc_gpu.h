// Device QLFC front end (see qlfc_front.hip).
#pragma once
#include "common.h"
#include <vector>

namespace qz {

class QlfcFront {
public:
    struct Seg {
        uint32_t block;   // block index in the batch
        uint32_t start;   // first BWT byte (global)
        uint32_t size;
        uint32_t r0, r1;  // runs [r0, r1)
    };
    ~QlfcFront() { release(); }
    int reserve(uint64_t n, uint32_t nblk);
    void release();
    // Splits every block of the batch (BWT bytes in device memory U, block
    // offsets offs) into coder segments and computes each segment's runs,
    // ranks and symbol table on the device.
    int run(const uint8_t* U, const std::vector<uint32_t>& offs, hipStream_t s);

    std::vector<Seg> segs_;
    uint32_t runs_ = 0;
    // device results: per run symbol / rank / length, per segment table
    uint8_t* rsym_ = nullptr;
    uint8_t* rrank_ = nullptr;
    uint32_t* rlen_ = nullptr;
    uint8_t* table_ = nullptr;   // segment x 256
    uint16_t* nsym_ = nullptr;   // per segment

private:
    uint32_t* flags_ = nullptr;
    uint32_t* scan_ = nullptr;
    uint32_t* tmp_ = nullptr;
    uint8_t* segflag_ = nullptr;
    void* sblk_ = nullptr;
    uint32_t* segstart_ = nullptr;
    uint32_t* rstart_ = nullptr;
    void* chunks_ = nullptr;
    uint32_t* segchunk_ = nullptr;
    uint8_t* sig_ = nullptr;
    uint16_t* siglen_ = nullptr;
    uint8_t* state_ = nullptr;
    uint16_t* statelen_ = nullptr;
    uint32_t* idx_ = nullptr;     // segment-start gather list
    uint32_t* gath_ = nullptr;
    uint64_t cap_ = 0;
    uint32_t cap_blk_ = 0;
    uint64_t cap_chunks_ = 0;
};

}  // namespace qz
