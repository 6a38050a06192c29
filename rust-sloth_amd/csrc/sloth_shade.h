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
    double* tri_acc;          // [n_tris][3 corners][6]: d color, d normal
    uint64_t n_slots;
    uint32_t n_tris, n_verts;
};

uint32_t shade_blocks(uint64_t n_slots);
int launch_shade(const ShadeArgs& a, bool backward, hipStream_t s);
int launch_shade_gather(const ShadeArgs& a, const uint32_t* adj_off, const uint32_t* adj, float* d_color,
                        float* d_normal, hipStream_t s);

}  // namespace sloth
