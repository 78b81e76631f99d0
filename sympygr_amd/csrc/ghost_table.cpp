// Host-only builder of the ghost refresh table (see ghost_table.h).
#include "ghost_table.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

namespace sgr {

// Ghost patch table of a block set (host only).
//
// Every block owns the half-open box of its interior nodes widened by half a
// cell.  Positions are taken to exact integers on the coarsest lattice of unit
// hmin / (2K), K = 1..64, on which every block origin and every half spacing
// is an integer, so ownership has no rounding ties.  For each block, every
// other block whose owned box meets its padded node box contributes one patch:
// the box of the target's nodes inside the source's owned box, clipped to
// the target's ghost region that is not beyond a flagged (domain-boundary)
// face.  The table is complete when these patches cover every ghost node of
// the target inside the domain exactly once, checked node by node on a
// per-target map; otherwise the reason is reported (nodes with no owner,
// nodes with two owners, overlapping blocks, no common lattice).

static int64_t ceil_div(int64_t a, int64_t b)  // b > 0
{
    return a >= 0 ? (a + b - 1) / b : -((-a) / b);
}

// Splits patch pi (class cls) into boxes for the ghost kernels.  Copy
// patches: at most GHOST_COPY_NODES nodes a box.  Interpolating patches:
// pieces n[d] per axis grown greedily until every box keeps the limits of
// k_ghost_sep (bounded with the largest box and the widest stencil span),
// each step taking the axis whose split costs least: the per-node loads
// NX*Ry*Rz + NY*ey*Rz + NZ*ey*ez summed over boxes, plus a fixed cost per box.
static void split_patch(const GhostPatch &g, uint32_t pi, int cls, std::vector<GhostBox> &out)
{
    const int nt[3] = {cls & 1 ? 1 : 6, cls & 2 ? 1 : 6, cls & 4 ? 1 : 6};
    int64_t n[3] = {1, 1, 1};
    auto side = [&](const int64_t *k, int d) { return (g.e[d] + k[d] - 1) / k[d]; };
    auto span = [&](int64_t s, int d) -> int64_t {  // upper bound of ghost_rows for s nodes
        return nt[d] == 1 ? s : ((s - 1) * g.ht[d] + g.hs[d] - 1) / g.hs[d] + 6;
    };
    auto fits = [&](const int64_t *k) {
        const int64_t s0 = side(k, 0), s1 = side(k, 1), s2 = side(k, 2);
        if (cls == GHOST_CLASSES - 1) return s0 * s1 * s2 <= GHOST_COPY_NODES;
        const int64_t Ry = span(s1, 1), Rz = span(s2, 2);
        return s0 <= GHOST_BOX_MAX && s1 <= GHOST_BOX_MAX && s2 <= GHOST_BOX_MAX && Ry <= GHOST_ROWS_MAX &&
               Rz <= GHOST_ROWS_MAX && s0 * Ry * Rz + s0 * s1 * Rz <= GHOST_LDS;
    };
    auto cost = [&](const int64_t *k) {
        const int64_t s0 = side(k, 0), s1 = side(k, 1), s2 = side(k, 2);
        const int64_t Ry = span(s1, 1), Rz = span(s2, 2);
        const int64_t per = nt[0] * s0 * Ry * Rz + nt[1] * s0 * s1 * Rz + nt[2] * s0 * s1 * s2;
        return k[0] * k[1] * k[2] * (per + 2048);
    };
    while (!fits(n)) {
        int best = -1;
        int64_t bc = 0;
        for (int d = 0; d < 3; d++) {
            if (side(n, d) <= 1) continue;
            int64_t m[3] = {n[0], n[1], n[2]};
            m[d]++;
            // copies: keep x rows whole while another axis can split
            const int64_t c = cls == GHOST_CLASSES - 1 ? (d == 0 ? 2 : 1) * (int64_t)1e12 - side(n, d) : cost(m);
            if (best < 0 || c < bc) best = d, bc = c;
        }
        n[best]++;  // all sides 1 always fits, so some side exceeds 1 here
    }
    for (int64_t c = 0; c < n[2]; c++)
        for (int64_t b = 0; b < n[1]; b++)
            for (int64_t a = 0; a < n[0]; a++) {
                const int64_t i[3] = {a, b, c};
                GhostBox x{};
                x.patch = (int32_t)pi;
                for (int d = 0; d < 3; d++) {
                    const int64_t l = g.e[d] * i[d] / n[d], h = g.e[d] * (i[d] + 1) / n[d];
                    x.lo[d] = g.lo[d] + (int32_t)l;
                    x.e[d] = (int32_t)(h - l);
                }
                if (x.e[0] > 0 && x.e[1] > 0 && x.e[2] > 0) out.push_back(x);
            }
}

GhostTable build_ghost_table(const sgr_block *blocks, uint32_t nb)
{
    GhostTable T;
    auto incomplete = [&](const std::string &m) {
        if (T.complete) T.complete = false, T.msg = m;
    };
    if (nb == 0) return T;
    struct Lat {
        int64_t P[3], H[3], m[3], lo[3], hi[3];
    };
    std::vector<Lat> L(nb);
    double O[3] = {HUGE_VAL, HUGE_VAL, HUGE_VAL}, hmin = HUGE_VAL;
    for (uint32_t b = 0; b < nb; b++)
        for (int d = 0; d < 3; d++) {
            O[d] = std::fmin(O[d], blocks[b].pmin[d]);
            hmin = std::fmin(hmin, (blocks[b].pmax[d] - blocks[b].pmin[d]) / (blocks[b].sz[d] - 1));
        }
    bool found = false;
    for (int K = 1; K <= 64 && !found; K++) {
        const double unit = hmin / (2.0 * K);
        found = true;
        for (uint32_t b = 0; b < nb && found; b++)
            for (int d = 0; d < 3 && found; d++) {
                const double h = (blocks[b].pmax[d] - blocks[b].pmin[d]) / (blocks[b].sz[d] - 1);
                const double a = (blocks[b].pmin[d] - O[d]) / unit, c = h / (2.0 * unit);
                if (std::fabs(a - std::nearbyint(a)) > 1e-6 || std::fabs(c - std::nearbyint(c)) > 1e-6 ||
                    std::fabs(a) > 1e9 || c > 1e9) {
                    found = false;
                    break;
                }
                Lat &l = L[b];
                l.P[d] = (int64_t)std::nearbyint(a);
                l.H[d] = 2 * (int64_t)std::nearbyint(c);
                l.m[d] = blocks[b].sz[d];
                l.lo[d] = l.P[d] + 5 * (l.H[d] / 2);
                l.hi[d] = l.P[d] + (l.m[d] - 4) * l.H[d] + l.H[d] / 2;
            }
    }
    if (!found) {
        incomplete("block coordinates share no common lattice (origins and half spacings must be integer "
                   "multiples of one unit)");
        return T;
    }
    for (uint32_t b = 0; b < nb; b++)
        for (int d = 0; d < 3; d++)
            if (L[b].P[d] + L[b].m[d] * L[b].H[d] >= (int64_t(1) << 30)) {
                incomplete("block set spans more than 2^30 lattice units");
                return T;
            }
    // candidate sources of a target: owned boxes sorted by their low x
    std::vector<uint32_t> order(nb);
    for (uint32_t b = 0; b < nb; b++) order[b] = b;
    std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t c) { return L[a].lo[0] < L[c].lo[0]; });
    int64_t wmax = 0;
    for (uint32_t b = 0; b < nb; b++) wmax = std::max(wmax, L[b].hi[0] - L[b].lo[0]);
    std::vector<GhostPatch> cls[GHOST_CLASSES];
    std::vector<uint8_t> owners;  // per node of the target's padded box: patches covering it
    for (uint32_t b = 0; b < nb && T.complete; b++) {
        const Lat &t = L[b];
        int64_t al[3], ah[3];
        for (int d = 0; d < 3; d++) {
            al[d] = (blocks[b].bflag >> (2 * d)) & 1 ? 3 : 0;
            ah[d] = (blocks[b].bflag >> (2 * d + 1)) & 1 ? t.m[d] - 3 : t.m[d];
        }
        owners.assign((size_t)(t.m[0] * t.m[1] * t.m[2]), 0);
        const int64_t x0 = t.P[0], x1 = t.P[0] + (t.m[0] - 1) * t.H[0];
        auto it = std::lower_bound(order.begin(), order.end(), x0 - wmax,
                                   [&](uint32_t s, int64_t v) { return L[s].lo[0] < v; });
        for (; it != order.end() && L[*it].lo[0] <= x1; ++it) {
            const uint32_t sidx = *it;
            if (sidx == b) continue;
            const Lat &s = L[sidx];
            int64_t lo[3], hi[3];
            bool empty = false;
            for (int d = 0; d < 3; d++) {
                lo[d] = std::max(al[d], ceil_div(s.lo[d] - t.P[d], t.H[d]));
                hi[d] = std::min(ah[d], ceil_div(s.hi[d] - t.P[d], t.H[d]));
                empty |= lo[d] >= hi[d];
            }
            if (empty) continue;
            bool meets_interior = true;
            for (int d = 0; d < 3; d++) meets_interior &= lo[d] < t.m[d] - 3 && hi[d] > 3;
            if (meets_interior) {
                incomplete("blocks " + std::to_string(b) + " and " + std::to_string(sidx) + " overlap");
                break;
            }
            GhostPatch g{};
            g.dst_off = blocks[b].offset;
            g.src_off = blocks[sidx].offset;
            g.dn0 = (int32_t)t.m[0];
            g.dn1 = (int32_t)t.m[1];
            g.sn0 = (int32_t)s.m[0];
            g.sn1 = (int32_t)s.m[1];
            int c = 0;
            uint64_t vol = 1;
            for (int d = 0; d < 3; d++) {
                g.lo[d] = (int32_t)lo[d];
                g.e[d] = (int32_t)(hi[d] - lo[d]);
                g.rel[d] = (int32_t)(t.P[d] - s.P[d]);
                g.ht[d] = (int32_t)t.H[d];
                g.hs[d] = (int32_t)s.H[d];
                g.sm[d] = (int32_t)s.m[d];
                g.inv_hs[d] = 1.0 / (double)s.H[d];
                if (t.H[d] % s.H[d] == 0 && (t.P[d] - s.P[d]) % s.H[d] == 0) c |= 1 << d;
                vol *= (uint64_t)g.e[d];
                // every tap an interior source node (owned box => q in [2, m-4])
                const int64_t lat0 = g.rel[d] + lo[d] * t.H[d], lat1 = g.rel[d] + (hi[d] - 1) * t.H[d];
                if (lat0 / s.H[d] < 2 || lat1 / s.H[d] > s.m[d] - 4 || s.m[d] < 12) {
                    incomplete("ghost patch of block " + std::to_string(b) + " leaves the interior of block " +
                               std::to_string(sidx));
                }
            }
            for (int64_t k = lo[2]; k < hi[2]; k++)
                for (int64_t j = lo[1]; j < hi[1]; j++)
                    for (int64_t i = lo[0]; i < hi[0]; i++) {
                        uint8_t &o = owners[(size_t)(i + t.m[0] * (j + t.m[1] * k))];
                        o = o < 2 ? o + 1 : 2;
                    }
            T.nodes += vol;
            if (c != GHOST_CLASSES - 1) T.interp_nodes += vol;
            cls[c].push_back(g);
        }
        // every ghost node inside the domain owned exactly once
        uint64_t none = 0, twice = 0;
        for (int64_t k = al[2]; k < ah[2] && T.complete; k++)
            for (int64_t j = al[1]; j < ah[1]; j++)
                for (int64_t i = al[0]; i < ah[0]; i++) {
                    if (i >= 3 && i < t.m[0] - 3 && j >= 3 && j < t.m[1] - 3 && k >= 3 && k < t.m[2] - 3) continue;
                    const uint8_t o = owners[(size_t)(i + t.m[0] * (j + t.m[1] * k))];
                    none += o == 0;
                    twice += o > 1;
                }
        if (T.complete && twice)
            incomplete("block " + std::to_string(b) + ": " + std::to_string(twice) +
                       " ghost nodes lie in the owned boxes of two blocks (overlapping blocks)");
        if (T.complete && none)
            incomplete("block " + std::to_string(b) + ": " + std::to_string(none) +
                       " ghost nodes inside the domain have no owning block (they lie outside every block)");
    }
    if (!T.complete) return T;
    for (int c = 0; c < GHOST_CLASSES; c++) {
        T.start[c] = (uint32_t)T.patches.size();
        T.bstart[c] = (uint32_t)T.boxes.size();
        for (const GhostPatch &g : cls[c]) {
            const size_t b0 = T.boxes.size();
            split_patch(g, (uint32_t)T.patches.size(), c, T.boxes);
            // every box within the limits k_ghost_sep / k_ghost_copy hold
            // (exact source rows), and the boxes tile the patch
            const int nt[3] = {c & 1 ? 1 : 6, c & 2 ? 1 : 6, c & 4 ? 1 : 6};
            uint64_t vol = 0;
            for (size_t k = b0; k < T.boxes.size(); k++) {
                const GhostBox &x = T.boxes[k];
                const uint64_t v = (uint64_t)x.e[0] * x.e[1] * x.e[2];
                vol += v;
                bool ok = true;
                for (int d = 0; d < 3; d++) ok &= x.lo[d] >= g.lo[d] && x.lo[d] + x.e[d] <= g.lo[d] + g.e[d];
                if (c == GHOST_CLASSES - 1) {
                    ok &= v <= (uint64_t)GHOST_COPY_NODES;
                } else {
                    const int Ry = ghost_rows(x.lo[1], x.e[1], g.rel[1], g.ht[1], g.hs[1], g.sm[1], nt[1]);
                    const int Rz = ghost_rows(x.lo[2], x.e[2], g.rel[2], g.ht[2], g.hs[2], g.sm[2], nt[2]);
                    ok &= x.e[0] <= GHOST_BOX_MAX && x.e[1] <= GHOST_BOX_MAX && x.e[2] <= GHOST_BOX_MAX &&
                          Ry <= GHOST_ROWS_MAX && Rz <= GHOST_ROWS_MAX &&
                          (int64_t)x.e[0] * Ry * Rz + (int64_t)x.e[0] * x.e[1] * Rz <= GHOST_LDS;
                }
                if (!ok) {
                    incomplete("internal error: a ghost box exceeds the kernel limits");
                    return T;
                }
            }
            if (vol != (uint64_t)g.e[0] * g.e[1] * g.e[2]) {
                incomplete("internal error: ghost boxes do not tile their patch");
                return T;
            }
            T.patches.push_back(g);
        }
    }
    T.start[GHOST_CLASSES] = (uint32_t)T.patches.size();
    T.bstart[GHOST_CLASSES] = (uint32_t)T.boxes.size();
    return T;
}

}  // namespace sgr
