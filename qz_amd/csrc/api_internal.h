// Device-batch helpers shared by the C-ABI translation units.
#pragma once
#include "pool.h"
#include <vector>

namespace qz {
int fwd_batch_device(Slot& sl, const uint8_t* d_in, uint8_t* d_out, const std::vector<uint32_t>& offs, int* primary,
                     int* aux, float* gpu_ms);
int inv_batch_device(Slot& sl, const uint8_t* d_in, uint8_t* d_out, const std::vector<uint32_t>& offs,
                     const int* primary, float* gpu_ms);
}  // namespace qz
