// Kernel entry points of libsloth_hip.so (internal).
#pragma once
#include "sloth_device.h"

namespace sloth {

// ---- general path (any width) -------------------------------------------
__global__ void k_clear_global(uint64_t* keys, uint64_t n_keys, int32_t* nl, uint64_t n_rows);
__global__ void k_raster_global(const Draw* draws, uint32_t d_lo, const Frame* frames, uint64_t* keys, int32_t* nl);
__global__ void k_winner_global(const Draw* draws, const Ramp* ramps, const Frame* frames, uint32_t f_lo,
                                uint64_t* keys, float* zbuf, int32_t* win, uint32_t win_raw);
__global__ void k_resolve_global(const Frame* frames, uint32_t f_lo, const uint64_t* keys, const int32_t* nl,
                                 uint32_t* cells);

}  // namespace sloth
