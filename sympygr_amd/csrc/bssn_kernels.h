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
    int derivs_threads;          // k_derivs workgroup size: 256 or 512
};

// Ghost refresh: one box of ghost nodes of a target block whose positions lie
// in the owned box of one source block (its interior nodes widened by half a
// cell).  Positions are integers on a lattice common to all blocks: target
// node i sits at rel + i*ht relative to source node 0, source nodes are hs
// apart.  Per axis a ghost node takes the coinciding source node or the
// 6-point Lagrange interpolant over source interior nodes.
struct GhostPatch {
    uint64_t dst_off, src_off;   // block offsets in the field arrays
    int32_t dn0, dn1;            // target padded x, y sizes
    int32_t sn0, sn1;            // source padded x, y sizes
    int32_t lo[3], e[3];         // target node box (padded coords) and extents
    int32_t rel[3];              // target node 0 - source node 0 (lattice units)
    int32_t ht[3], hs[3];        // spacings (lattice units)
    int32_t sm[3];               // source padded sizes
    double inv_hs[3];            // 1.0 / hs
};

// Patch classes: bit d set when every target node of the box coincides with
// a source node along axis d (one tap instead of six).  Patches are stored
// sorted by class; class c occupies [start[c], start[c+1]).
constexpr int GHOST_CLASSES = 8;

// Unit of work of the ghost kernels (one workgroup): the target nodes
// lo .. lo+e-1 (padded coordinates of the target block) of patches[patch].
// Boxes are stored sorted by their patch's class, like the patches.
struct GhostBox {
    int32_t patch;
    int32_t lo[3], e[3];
    int32_t pad;
};

// Limits of an interpolating box (classes 0..6), kept by the host when it
// splits patches: extent per axis, source rows held per axis for y and z,
// and the LDS doubles of the two partial-sum arrays, ex*Ry*Rz + ex*ey*Rz.
constexpr int GHOST_BOX_MAX = 64;
constexpr int GHOST_ROWS_MAX = 128;
constexpr int GHOST_LDS = 4096;
// Nodes per box of the copy class (7).
constexpr int GHOST_COPY_NODES = 2048;

// First source node of the taps of target node idx along one axis: the
// coinciding node (nt == 1) or the start of the 6-point stencil, kept on
// source nodes 3 .. sm-4 (interior).  Target positions lie inside the
// source's owned box, so lat > 0.
__host__ __device__ inline int ghost_p0(int idx, int rel, int ht, int hs, int sm, int nt)
{
    const int q = (rel + idx * ht) / hs;
    if (nt == 1) return q;
    const int p = q - 2 < 3 ? 3 : q - 2;
    return p < sm - 9 ? p : sm - 9;
}

// Source rows an axis of a box needs: the span of its stencils (nt == 6),
// or one row per target node (nt == 1).
__host__ __device__ inline int ghost_rows(int lo, int e, int rel, int ht, int hs, int sm, int nt)
{
    if (nt == 1) return e;
    return ghost_p0(lo + e - 1, rel, ht, hs, sm, 6) - ghost_p0(lo, rel, ht, hs, sm, 6) + 6;
}

// Hamiltonian (cout[0]) and momentum (cout[1..3]) constraints; uses the workspace.
hipError_t launch_constraints(const LaunchArgs &a, OutPtrs cout, hipStream_t s);

// boxes sorted by class: class c occupies [bstart[c], bstart[c+1])
hipError_t launch_ghost_fill(OutPtrs u, const GhostPatch *patches, const GhostBox *boxes,
                             const uint32_t bstart[GHOST_CLASSES + 1], hipStream_t s);

hipError_t launch_rhs(const LaunchArgs &a, hipStream_t s);    // derivative workspace + pointwise
hipError_t launch_fused(const LaunchArgs &a, hipStream_t s);  // fused single pass

// tile shapes (interior points) of the two kernel variants
constexpr int FUSED_TX = 8, FUSED_TY = 8, FUSED_TZ = 4;

}  // namespace sgr
