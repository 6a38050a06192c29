// Tiled raster path (even widths): declared here, implemented in sloth_tiled.hip.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "sloth_device.h"

namespace sloth {

struct TiledPlan {
    int plan(const std::vector<Frame>&, const std::vector<Draw>&, uint64_t) { return 5; }
    int upload(hipStream_t) { return 5; }
    int launch(const std::vector<Frame>&, const std::vector<Draw>&, const Frame*, const Draw*, const Ramp*, uint32_t,
               uint32_t, uint32_t*, float*, int32_t*, hipStream_t, std::vector<hipEvent_t>*)
    {
        return 5;
    }
    static std::vector<const char*> stage_names() { return {}; }
};

}  // namespace sloth
