// Kernel launch interface used by the C-ABI layer.
#pragma once

#include "bssn_common.h"

namespace sgr {

// Fused classic-RK4 stage update applied at the output store (stage >= 0):
//   stage 0: acc  = u0 + w k            next = u0 + cn k
//   stage 1,2: acc = acc + w k          next = u0 + cn k
//   stage 3: next = acc + w k           (the new state; may alias u0)
// with (cn, w) = (dt/2, dt/6), (dt/2, dt/3), (dt, dt/3), (-, dt/6).
struct RkArgs {
    int stage;  // -1: plain RHS into `out`
    double cn, w;
    FieldPtrs u0;
    OutPtrs acc;
    OutPtrs next;
};

struct LaunchArgs {
    FieldPtrs in;
    OutPtrs out;
    const BlockDesc *blocks;     // device
    const uint32_t *tile_block;  // device, block id of every tile
    uint32_t ntiles;             // tiles of this launch (one workspace chunk)
    uint32_t tile0;              // first tile of this launch
    void *ws;                    // derivative workspace [slot][np padded to the stride] (double or float)
    double *wsK;                 // KO sums [24][stride]
    uint64_t np;                 // plane stride in points (>= ntiles * 512)
    Params params;
    bool upwind;
    bool mixed;
    hipEvent_t ev[4];            // optional: start/stop of k_derivs, start/stop of k_pointwise
    RkArgs rk;                   // rk.stage < 0: plain RHS
};

// Same-level ghost refresh: copy one box of ghost points of a block from the
// interior of its neighbour (all 24 fields).
struct GhostCopy {
    uint64_t dst_off, src_off;   // block offsets in the field arrays
    int32_t n0, n1;              // padded x, y sizes (shared by both blocks)
    int32_t d[3];                // first destination point (padded coords)
    int32_t s[3];                // first source point
    int32_t e[3];                // extents
};

// Hamiltonian (cout[0]) and momentum (cout[1..3]) constraints; uses the workspace.
hipError_t launch_constraints(const LaunchArgs &a, OutPtrs cout, hipStream_t s);

hipError_t launch_ghost_fill(OutPtrs u, const GhostCopy *copies, uint32_t ncopies, hipStream_t s);

hipError_t launch_rhs(const LaunchArgs &a, hipStream_t s);    // derivative workspace + pointwise
hipError_t launch_fused(const LaunchArgs &a, hipStream_t s);  // fused single pass

// tile shapes (interior points) of the two kernel variants
constexpr int FUSED_TX = 8, FUSED_TY = 8, FUSED_TZ = 4;

}  // namespace sgr
