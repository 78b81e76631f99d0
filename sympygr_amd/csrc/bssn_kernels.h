// Kernel launch interface used by the C-ABI layer.
#pragma once

#include "bssn_common.h"
#include "ghost_table.h"

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
    uint32_t nfast;              // RHS: the first nfast tiles of the launch have the 8^3 shape, the rest others
    uint32_t *queue;             // RHS: 8 tile-queue counters (device; null: one stencil workgroup per tile)
    uint32_t grid_cap;           // RHS: resident stencil workgroups (2 per CU), a multiple of 8
    void *ws;                    // derivative workspace (double or float), tile-major
    double *wsT;                 // RHS pass: tails of the pointwise outputs (one per field and point)
    double *wsK;                 // constraint pass: KO sums (24 per point)
    Params params;
    bool upwind;
    bool mixed;
    bool ccz4;                   // the CCZ4 system (25 fields) instead of BSSN
    hipEvent_t ev[4];            // optional: start/stop of k_stencil, start/stop of k_pointwise
    RkArgs rk;                   // rk.stage < 0: plain RHS
    int derivs_threads;          // constraint pass: k_derivs workgroup size, 256 or 512
    // RHS dissipation: 0 Kreiss-Oliger (in the stencil pass, params.ko_sigma),
    // 1 TVB3, 2 TVB5 (k_tvb adds diss_sigma times it to the tails; the
    // stencil pass then runs with params.ko_sigma = 0)
    int diss;
    double diss_sigma;
    double *lam;                 // TVB: speeds along x, y, z, lam_span nodes each
    uint64_t lam_span;
    int only;                    // RHS: 0 both passes, 1 the stencil pass only, 2 the rest only
};

// Hamiltonian (cout[0]) and momentum (cout[1..3]) constraints, with psi4 also
// Re/Im Psi4 (cout[4], cout[5]); uses the workspace.
hipError_t launch_constraints(const LaunchArgs &a, OutPtrs cout, bool psi4, hipStream_t s);

// boxes sorted by class: class c occupies [bstart[c], bstart[c+1])
hipError_t launch_ghost_fill(OutPtrs u, int nf, const GhostPatch *patches, const GhostBox *boxes,
                             const uint32_t bstart[GHOST_CLASSES + 1], hipStream_t s);

hipError_t launch_rhs(const LaunchArgs &a, hipStream_t s);    // derivative workspace + pointwise

// TVB dissipation: the characteristic speeds at every node (before launch_rhs)
hipError_t launch_speeds(const LaunchArgs &a, hipStream_t s);

}  // namespace sgr
