// Kernel launch interface used by the C-ABI layer.
#pragma once

#include "bssn_common.h"

namespace sgr {

struct LaunchArgs {
    FieldPtrs in;
    OutPtrs out;
    const BlockDesc *blocks;     // device
    const uint32_t *tile_block;  // device, block id of every tile
    uint32_t ntiles;
    void *ws;                    // derivative workspace [slot][np padded to the stride] (double or float)
    double *wsK;                 // KO sums [24][stride]
    uint64_t np;                 // plane stride in points (>= ntiles * 512)
    Params params;
    bool upwind;
    bool mixed;
    hipEvent_t ev[4];            // optional: start/stop of k_derivs, start/stop of k_pointwise
};

hipError_t launch_rhs(const LaunchArgs &a, hipStream_t s);    // derivative workspace + pointwise
hipError_t launch_fused(const LaunchArgs &a, hipStream_t s);  // fused single pass

// tile shapes (interior points) of the two kernel variants
constexpr int FUSED_TX = 8, FUSED_TY = 8, FUSED_TZ = 4;

}  // namespace sgr
