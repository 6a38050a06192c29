// Differentiable shading kernels (internal); see sloth_shade.hip.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>

#include "sloth_device.h"
#include "sloth_host.h"

namespace sloth {

struct ShadeArgs {
    const Frame* frames;
    uint32_t n_frames;
    const Draw* draws;
    const int32_t* vis;       // per z slot: (triangle << 1 | sub), -1 = none
    const uint32_t* faces;    // [n_tris][3] vertex ids (mesh-local triangle index)
    const float* vcolor;      // [n_verts][3]
    const float* vnormal;     // [n_verts][3]
    const float* light;       // [3]
    const float* G;           // per draw, 3x3 row-major normal map
    const float* target;      // per z slot [3], or null (no loss)
    float* out;               // per z slot [3], or null
    double* partial;          // per block: loss, dL.x, dL.y, dL.z
    double* vert_acc;         // [n_verts][6]: d color, d normal (binary64 sums)
    uint64_t n_slots;
    uint32_t n_tris, n_verts;
    const uint64_t* bwd_tile_begin;   // per frame (+1): first backward tile
    uint64_t n_bwd_tiles;
};

uint32_t shade_blocks(uint64_t n_slots);
// workgroups (and per-block partial sums) of the backward pass
uint32_t bwd_blocks(uint64_t n_slots, uint64_t n_bwd_tiles);
// Backward tiles per frame, prefix-summed: begin[0..n_frames]
void bwd_tiles(const Frame* frames, uint32_t n_frames, uint64_t* begin);
int launch_shade(const ShadeArgs& a, bool backward, hipStream_t s);
// binary64 vertex sums -> binary32 gradients
int launch_shade_finish(const ShadeArgs& a, float* d_color, float* d_normal, hipStream_t s);

}  // namespace sloth
