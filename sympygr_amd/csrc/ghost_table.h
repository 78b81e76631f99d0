// Ghost refresh tables: the patch and box types the ghost kernels read, and
// the host-only builder of a block set's table.  This header and
// ghost_table.cpp need no HIP compiler, so the builder also compiles with a
// plain C++ compiler (for instance under the address and undefined-behaviour
// sanitizers).
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/sympygr_amd.h"

#if !defined(__HIPCC__) && !defined(__host__)
#define __host__
#define __device__
#endif

namespace sgr {

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

// A block set's ghost refresh table: patches and their kernel work units.
struct GhostTable {
    std::vector<GhostPatch> patches;  // sorted by class
    uint32_t start[GHOST_CLASSES + 1] = {};
    std::vector<GhostBox> boxes;      // patches split into kernel work units, sorted by class
    uint32_t bstart[GHOST_CLASSES + 1] = {};
    uint64_t nodes = 0, interp_nodes = 0;
    bool complete = true;
    std::string msg;
};

// Table of `nb` blocks whose sizes are at least 13 per axis (callers
// validate); `complete` is false, with the reason in `msg`, when the refresh
// would be refused.
GhostTable build_ghost_table(const sgr_block *blocks, uint32_t nb);

}  // namespace sgr
