// Output encoders (internal); see sloth_encode.hip.
#pragma once
#include <hip/hip_runtime.h>

#include "sloth_host.h"

namespace sloth {

uint64_t encode_blocks(uint64_t n_cells);
// lengths -> exclusive offsets per cell (off), per scan block (block_sum), and
// the total byte count (*total, device memory)
int launch_encode_size(const uint32_t* cells, uint64_t n, int mode, uint64_t* off, uint64_t* block_sum,
                       uint64_t* total, hipStream_t s);
int launch_encode_write(const uint32_t* cells, uint64_t n, int mode, const uint64_t* off, const uint64_t* block_sum,
                        uint8_t* out, hipStream_t s);

}  // namespace sloth
