// C ABI of libsympygr_amd.so: plans (validated, uploaded block tables plus a
// workspace allocated once), the launch entry points and error reporting.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <new>
#include <string>
#include <vector>

#include "../../include/sympygr_amd.h"
#include "bssn_kernels.h"

using namespace sgr;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

int hip_fail(hipError_t e, const char *what)
{
    return fail(SGR_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(call, what)                                                                                   \
    do {                                                                                                      \
        hipError_t e_ = (call);                                                                               \
        if (e_ != hipSuccess) return hip_fail(e_, what);                                                      \
    } while (0)

// Stencil constants with the reference's expressions (bssn/src/derivs.cpp:
// 2058-2061, 2446-2449, 499-502, 919-923); h as bssn/src/rhs.cpp:144-146.
AxisConst axis_const(double pmin, double pmax, uint32_t n)
{
    AxisConst a;
    a.h = (pmax - pmin) / (n - 1);
    const double idx = 1.0 / a.h;
    a.by12 = idx / 12.0;
    a.by60 = idx / 60.0;
    const double idx_sqrd = 1.0 / (a.h * a.h);
    a.sq180 = idx_sqrd / 180.0;
    a.sq12 = idx_sqrd / 12.0;
    a.by2 = 0.50 * idx;
    a.ko_pre = -1.0 / 64.0 / a.h;
    a.smr3 = 59.0 / 48.0 * 64 * a.h;
    a.smr2 = 43.0 / 48.0 * 64 * a.h;
    a.smr1 = 49.0 / 48.0 * 64 * a.h;
    return a;
}

int validate_params(const sgr_params *p)
{
    if (!p) return fail(SGR_E_INVALID, "params is null");
    if (p->adv_mode != SGR_ADV_CENTRED && p->adv_mode != SGR_ADV_UPWIND)
        return fail(SGR_E_INVALID, "adv_mode must be SGR_ADV_CENTRED or SGR_ADV_UPWIND");
    if (p->precision != SGR_PREC_FP64 && p->precision != SGR_PREC_MIXED)
        return fail(SGR_E_INVALID, "precision must be SGR_PREC_FP64 or SGR_PREC_MIXED");
    if (p->gauge != SGR_GAUGE_STANDARD && p->gauge != SGR_GAUGE_ROCHESTER)
        return fail(SGR_E_INVALID, "gauge must be SGR_GAUGE_STANDARD or SGR_GAUGE_ROCHESTER");
    if (p->eta_mode != SGR_ETA_CONST && p->eta_mode != SGR_ETA_FUNC)
        return fail(SGR_E_INVALID, "eta_mode must be SGR_ETA_CONST or SGR_ETA_FUNC");
    return SGR_OK;
}

int validate_block(const sgr_block &b, uint32_t idx)
{
    char msg[256];
    for (int d = 0; d < 3; d++) {
        if (b.sz[d] < 13) {
            snprintf(msg, sizeof msg, "block %u: sz[%d]=%u < 13 (stencils need 7 interior points)", idx, d, b.sz[d]);
            return fail(SGR_E_INVALID, msg);
        }
        if (!(b.pmax[d] > b.pmin[d]) || !std::isfinite(b.pmin[d]) || !std::isfinite(b.pmax[d])) {
            snprintf(msg, sizeof msg, "block %u: pmax[%d] must be finite and > pmin[%d]", idx, d, d);
            return fail(SGR_E_INVALID, msg);
        }
    }
    if (b.bflag > 63) {
        snprintf(msg, sizeof msg, "block %u: bflag %u has bits above the six faces", idx, b.bflag);
        return fail(SGR_E_INVALID, msg);
    }
    return SGR_OK;
}

Params device_params(const sgr_params &p)
{
    Params q;
    for (int i = 0; i < 4; i++) q.lambda[i] = (double)p.lambda[i];
    q.lambda_f[0] = p.lambda_f[0];
    q.lambda_f[1] = p.lambda_f[1];
    q.eta_const = p.eta_const;
    q.eta_r0 = p.eta_r0;
    q.eta_damping_exp = p.eta_damping_exp;
    q.ko_sigma = p.ko_sigma;
    q.rochester = p.gauge == SGR_GAUGE_ROCHESTER;
    q.eta_func = p.eta_mode == SGR_ETA_FUNC;
    for (int i = 0; i < 3; i++) q.xi[i] = (double)p.xi[i];
    q.eta_func_r0 = p.eta_func_r0;
    q.eta_power[0] = p.eta_power[0];
    q.eta_power[1] = p.eta_power[1];
    return q;
}

__global__ void k_check_finite(FieldPtrs in, const BlockDesc *blocks, uint32_t nblocks, int *flag)
{
    for (uint32_t b = blockIdx.y; b < nblocks; b += gridDim.y) {
        const BlockDesc &B = blocks[b];
        const uint64_t n = (uint64_t)B.n[0] * B.n[1] * B.n[2];
        for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < n; e += (uint64_t)gridDim.x * blockDim.x) {
            bool bad = false;
            for (int f = 0; f < NF; f++) bad |= !isfinite(in.p[f][B.offset + e]);
            if (bad) atomicOr(flag, 1);
        }
    }
}

// Plain HBM copy, 16 bytes per lane, grid-stride (a reference point of known
// traffic for hardware byte counters).
__global__ __launch_bounds__(256) void k_hbm_copy(double2 *__restrict__ dst, const double2 *__restrict__ src, uint64_t n)
{
    for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256ull) dst[i] = src[i];
}

}  // namespace

struct GhostTable {
    std::vector<GhostPatch> patches;  // sorted by class
    uint32_t start[GHOST_CLASSES + 1] = {};
    std::vector<GhostBox> boxes;      // patches split into kernel work units, sorted by class
    uint32_t bstart[GHOST_CLASSES + 1] = {};
    uint64_t nodes = 0, interp_nodes = 0;
    bool complete = true;
    std::string msg;
};

enum class Variant { Fused, Workspace };

struct sgr_plan {
    Variant variant = Variant::Workspace;
    int derivs_threads = 512;  // k_derivs workgroup size (512: one point per thread, 256: two)
    int device = 0;
    sgr_params params{};
    std::vector<BlockDesc> hblocks;
    BlockDesc *d_blocks = nullptr;
    uint32_t *d_tile_block = nullptr;
    uint32_t ntiles = 0;
    void *d_ws = nullptr;
    double *d_wsK = nullptr;
    int *d_flag = nullptr;
    uint64_t np = 0;             // workspace plane stride (points)
    uint32_t chunk_tiles = 0;    // tiles per launch (workspace capacity)
    uint64_t ws_bytes = 0;
    uint64_t interior = 0;
    uint32_t nslots = 0;
    // ghost refresh table
    GhostTable ghost;
    GhostPatch *d_ghost = nullptr;
    GhostBox *d_gbox = nullptr;
};

// Ghost patch table of a block set (host only).
//
// Every block owns the half-open box of its interior nodes widened by half a
// cell.  Positions are taken to exact integers on the coarsest lattice of unit
// hmin / (2K), K = 1..64, on which every block origin and every half spacing
// is an integer, so ownership has no rounding ties.  For each block, every
// other block whose owned box meets its padded node box contributes one patch:
// the box of the target's nodes inside the source's owned box, clipped to
// the target's ghost region that is not beyond a flagged (domain-boundary)
// face.  The table is complete when these patches cover exactly the target's
// ghost nodes inside the domain; otherwise the reason is reported (nodes with
// no owner, overlapping blocks, no common lattice).

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

static GhostTable build_ghost_table(const sgr_block *blocks, uint32_t nb)
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
    for (uint32_t b = 0; b < nb && T.complete; b++) {
        const Lat &t = L[b];
        int64_t al[3], ah[3];
        uint64_t required = 1, interior = 1, covered = 0;
        for (int d = 0; d < 3; d++) {
            al[d] = (blocks[b].bflag >> (2 * d)) & 1 ? 3 : 0;
            ah[d] = (blocks[b].bflag >> (2 * d + 1)) & 1 ? t.m[d] - 3 : t.m[d];
            required *= (uint64_t)(ah[d] - al[d]);
            interior *= (uint64_t)(t.m[d] - 6);
        }
        required -= interior;
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
            covered += vol;
            T.nodes += vol;
            if (c != GHOST_CLASSES - 1) T.interp_nodes += vol;
            cls[c].push_back(g);
        }
        if (T.complete && covered != required)
            incomplete("block " + std::to_string(b) + ": " + std::to_string(covered) + " of " +
                       std::to_string(required) + " ghost nodes inside the domain have an owning block" +
                       (covered > required ? " (overlapping blocks)" : " (the rest lie outside every block)"));
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

static void plan_free(sgr_plan *p)
{
    if (!p) return;
    int cur = 0;
    hipGetDevice(&cur);
    hipSetDevice(p->device);
    hipFree(p->d_blocks);
    hipFree(p->d_tile_block);
    hipFree(p->d_ws);
    hipFree(p->d_wsK);
    hipFree(p->d_flag);
    hipFree(p->d_ghost);
    hipFree(p->d_gbox);
    hipSetDevice(cur);
    delete p;
}

extern "C" {

void sgr_params_default(sgr_params *p, int32_t adv_mode, int32_t precision)
{
    if (!p) return;
    for (int i = 0; i < 4; i++) p->lambda[i] = 1;
    p->lambda_f[0] = 1.0;
    p->lambda_f[1] = 0.0;
    p->eta_const = 2.0;
    p->eta_r0 = 50.0;
    p->eta_damping_exp = 1.0;
    p->ko_sigma = 0.01;
    p->adv_mode = adv_mode;
    p->precision = precision;
    p->gauge = SGR_GAUGE_STANDARD;
    p->eta_mode = SGR_ETA_CONST;
    for (int i = 0; i < 3; i++) p->xi[i] = 0;
    p->eta_func_r0 = 1.31;
    p->eta_power[0] = 2.0;
    p->eta_power[1] = 2.0;
}

int sgr_plan_create(sgr_plan **plan, const sgr_block *blocks, uint32_t nblocks, const sgr_params *params, int device)
{
    if (!plan) return fail(SGR_E_INVALID, "plan is null");
    *plan = nullptr;
    if (nblocks > 0 && !blocks) return fail(SGR_E_INVALID, "blocks is null");
    int rc = validate_params(params);
    if (rc) return rc;
    sgr_plan *p = new (std::nothrow) sgr_plan;
    if (!p) return fail(SGR_E_NOMEM, "host allocation failed");
    p->device = device;
    p->params = *params;
    // Default: the two-pass kernels (LDS-tiled derivative kernel into an HBM
    // workspace, then two output-group pointwise kernels).  SGR_KERNEL=fused
    // selects the single-pass kernel (no workspace; slower on gfx950 today
    // because its register footprint allows one wave per SIMD).
    // SGR_DERIVS_THREADS=256|512 sets the derivative kernel's workgroup size.
    {
        const char *kv = getenv("SGR_KERNEL");
        if (kv && !strcmp(kv, "fused")) p->variant = Variant::Fused;
        const char *tv = getenv("SGR_DERIVS_THREADS");
        if (tv && !strcmp(tv, "512")) p->derivs_threads = 512;
        else if (tv && !strcmp(tv, "256")) p->derivs_threads = 256;
    }
    const bool fusedv = p->variant == Variant::Fused;
    if (fusedv && (params->gauge != SGR_GAUGE_STANDARD || params->eta_mode != SGR_ETA_CONST)) {
        delete p;
        return fail(SGR_E_UNSUPPORTED, "the fused kernel evaluates the standard gauge with constant eta only");
    }
    const int tdim[3] = {fusedv ? FUSED_TX : T, fusedv ? FUSED_TY : T, fusedv ? FUSED_TZ : T};
    p->nslots = fusedv ? 0 : NSLOT;
    std::vector<uint32_t> tile_block;
    uint64_t ws_base = 0;
    for (uint32_t b = 0; b < nblocks; b++) {
        rc = validate_block(blocks[b], b);
        if (rc) {
            delete p;
            return rc;
        }
        BlockDesc d{};
        d.offset = blocks[b].offset;
        d.ws_base = ws_base;
        d.bflag = blocks[b].bflag;
        d.tile_base = (uint32_t)tile_block.size();
        uint64_t nt = 1;
        for (int k = 0; k < 3; k++) {
            d.n[k] = (int32_t)blocks[b].sz[k];
            d.ntile[k] = (int32_t)((blocks[b].sz[k] - 6 + tdim[k] - 1) / tdim[k]);
            d.pmin[k] = blocks[b].pmin[k];
            d.ax[k] = axis_const(blocks[b].pmin[k], blocks[b].pmax[k], blocks[b].sz[k]);
            nt *= (uint64_t)d.ntile[k];
        }
        if (tile_block.size() + nt > 0xFFFFFFF0u) {
            delete p;
            return fail(SGR_E_INVALID, "too many tiles");
        }
        for (uint64_t t = 0; t < nt; t++) tile_block.push_back(b);
        if (!fusedv) ws_base += nt * TILE_PTS;
        p->interior += (uint64_t)(blocks[b].sz[0] - 6) * (blocks[b].sz[1] - 6) * (blocks[b].sz[2] - 6);
        p->hblocks.push_back(d);
    }
    int ndev = 0;
    hipError_t derr = hipGetDeviceCount(&ndev);
    if (derr != hipSuccess || device < 0 || device >= ndev) {
        delete p;
        return derr != hipSuccess ? hip_fail(derr, "hipGetDeviceCount") : fail(SGR_E_INVALID, "device index out of range");
    }
    p->ntiles = (uint32_t)tile_block.size();
    // The workspace holds at most `chunk_tiles` tiles (budget SGR_WS_GIB,
    // default 40 GiB); larger plans run as several launches over equal tile
    // ranges.  Plane stride: the chunk's point count padded by a few KiB so
    // consecutive slot planes do not share one power-of-two alignment.
    const size_t wsz = params->precision == SGR_PREC_MIXED ? sizeof(float) : sizeof(double);
    {
        double gib = 40.0;
        const char *gv = getenv("SGR_WS_GIB");
        if (gv) gib = atof(gv);
        const uint64_t per_tile = (uint64_t)TILE_PTS * ((uint64_t)p->nslots * wsz + NF * sizeof(double));
        uint64_t cap = (uint64_t)(gib * 1073741824.0) / (per_tile ? per_tile : 1);
        if (cap < 1) cap = 1;
        const uint64_t nt = tile_block.size();
        // equal chunks (no short last launch)
        const uint64_t nchunks = nt ? (nt + cap - 1) / cap : 1;
        p->chunk_tiles = (uint32_t)(fusedv ? nt : (nt + nchunks - 1) / nchunks);
        uint64_t pad = 2048 + 64;
        const char *pv = getenv("SGR_WS_PAD");
        if (pv) pad = strtoull(pv, nullptr, 10);
        p->np = fusedv ? 0 : (uint64_t)p->chunk_tiles * TILE_PTS + (nt ? pad : 0);
    }
    p->ws_bytes = fusedv ? 0 : p->np * (uint64_t)p->nslots * wsz + p->np * (uint64_t)NF * sizeof(double);

    int cur = 0;
    hipGetDevice(&cur);
    hipError_t e = hipSetDevice(device);
    auto bail = [&](hipError_t err, const char *what) {
        hipSetDevice(cur);
        plan_free(p);
        return err == hipErrorOutOfMemory ? fail(SGR_E_NOMEM, std::string(what) + ": out of device memory")
                                          : hip_fail(err, what);
    };
    if (e != hipSuccess) return bail(e, "hipSetDevice");
    if (nblocks) {
        if ((e = hipMalloc(&p->d_blocks, nblocks * sizeof(BlockDesc))) != hipSuccess) return bail(e, "hipMalloc(blocks)");
        if ((e = hipMemcpy(p->d_blocks, p->hblocks.data(), nblocks * sizeof(BlockDesc), hipMemcpyHostToDevice)) != hipSuccess)
            return bail(e, "hipMemcpy(blocks)");
        if ((e = hipMalloc(&p->d_tile_block, tile_block.size() * sizeof(uint32_t))) != hipSuccess)
            return bail(e, "hipMalloc(tiles)");
        if ((e = hipMemcpy(p->d_tile_block, tile_block.data(), tile_block.size() * sizeof(uint32_t),
                           hipMemcpyHostToDevice)) != hipSuccess)
            return bail(e, "hipMemcpy(tiles)");
        if (!fusedv) {
            if ((e = hipMalloc(&p->d_ws, p->np * (uint64_t)p->nslots * wsz)) != hipSuccess)
                return bail(e, "hipMalloc(workspace)");
            if ((e = hipMalloc(&p->d_wsK, p->np * (uint64_t)NF * sizeof(double))) != hipSuccess)
                return bail(e, "hipMalloc(ko workspace)");
        }
    }
    if ((e = hipMalloc(&p->d_flag, sizeof(int))) != hipSuccess) return bail(e, "hipMalloc(flag)");
    p->ghost = build_ghost_table(blocks, nblocks);
    if (!p->ghost.patches.empty()) {
        const size_t gb = p->ghost.patches.size() * sizeof(GhostPatch);
        if ((e = hipMalloc(&p->d_ghost, gb)) != hipSuccess) return bail(e, "hipMalloc(ghost table)");
        if ((e = hipMemcpy(p->d_ghost, p->ghost.patches.data(), gb, hipMemcpyHostToDevice)) != hipSuccess)
            return bail(e, "hipMemcpy(ghost table)");
        const size_t bb = p->ghost.boxes.size() * sizeof(GhostBox);
        if ((e = hipMalloc(&p->d_gbox, bb)) != hipSuccess) return bail(e, "hipMalloc(ghost boxes)");
        if ((e = hipMemcpy(p->d_gbox, p->ghost.boxes.data(), bb, hipMemcpyHostToDevice)) != hipSuccess)
            return bail(e, "hipMemcpy(ghost boxes)");
    }
    hipSetDevice(cur);
    *plan = p;
    return SGR_OK;
}

int sgr_plan_destroy(sgr_plan *plan)
{
    plan_free(plan);
    return SGR_OK;
}

int sgr_plan_get_info(const sgr_plan *plan, sgr_plan_info *info)
{
    if (!plan || !info) return fail(SGR_E_INVALID, "null argument");
    info->blocks = plan->hblocks.size();
    info->tiles = plan->ntiles;
    info->interior_points = plan->interior;
    info->workspace_bytes = plan->ws_bytes;
    info->derivative_slots = plan->nslots;
    info->chunks = plan->chunk_tiles ? (plan->ntiles + plan->chunk_tiles - 1) / plan->chunk_tiles : 0;
    return SGR_OK;
}

static int rhs_impl(sgr_plan *plan, double *const out[SGR_NUM_VARS], const double *const in[SGR_NUM_VARS],
                    hipStream_t stream, hipEvent_t *ev, const RkArgs *rk = nullptr)
{
    if (!plan || !out || !in) return fail(SGR_E_INVALID, "null argument");
    LaunchArgs a{};
    for (int f = 0; f < NF; f++) {
        if (!in[f] || !out[f]) return fail(SGR_E_INVALID, "null field pointer");
        a.in.p[f] = in[f];
        a.out.p[f] = out[f];
    }
    a.blocks = plan->d_blocks;
    a.tile_block = plan->d_tile_block;
    a.ntiles = plan->ntiles;
    a.ws = plan->d_ws;
    a.wsK = plan->d_wsK;
    a.np = plan->np;
    a.params = device_params(plan->params);
    a.upwind = plan->params.adv_mode == SGR_ADV_UPWIND;
    a.mixed = plan->params.precision == SGR_PREC_MIXED;
    a.derivs_threads = plan->derivs_threads;
    for (int q = 0; q < 4; q++) a.ev[q] = ev ? ev[q] : nullptr;
    a.rk.stage = -1;
    if (rk) {
        if (plan->variant == Variant::Fused)
            return fail(SGR_E_UNSUPPORTED, "RK4 stages need the two-pass kernels (unset SGR_KERNEL=fused)");
        a.rk = *rk;
    }
    int cur = 0;
    hipGetDevice(&cur);
    if (cur != plan->device) HIP_TRY(hipSetDevice(plan->device), "hipSetDevice");
    hipError_t e = hipSuccess;
    if (plan->variant == Variant::Fused) {
        e = launch_fused(a, stream);
        // one kernel: the second event pair brackets nothing
        if (e == hipSuccess && ev) e = hipEventRecord(ev[2], stream);
        if (e == hipSuccess && ev) e = hipEventRecord(ev[3], stream);
    } else if (plan->ntiles <= plan->chunk_tiles) {
        e = launch_rhs(a, stream);
    } else {
        // several workspace chunks: events[0..1] bracket the whole call,
        // events[2..3] are recorded at its end
        if (ev) e = hipEventRecord(ev[0], stream);
        for (int q = 0; q < 4; q++) a.ev[q] = nullptr;
        for (uint32_t t0 = 0; e == hipSuccess && t0 < plan->ntiles; t0 += plan->chunk_tiles) {
            a.tile0 = t0;
            a.ntiles = plan->ntiles - t0 < plan->chunk_tiles ? plan->ntiles - t0 : plan->chunk_tiles;
            e = launch_rhs(a, stream);
        }
        if (e == hipSuccess && ev) e = hipEventRecord(ev[1], stream);
        if (e == hipSuccess && ev) e = hipEventRecord(ev[2], stream);
        if (e == hipSuccess && ev) e = hipEventRecord(ev[3], stream);
    }
    if (cur != plan->device) hipSetDevice(cur);
    if (e != hipSuccess) return hip_fail(e, "kernel launch");
    return SGR_OK;
}

int sgr_rhs(sgr_plan *plan, double *const out[SGR_NUM_VARS], const double *const in[SGR_NUM_VARS], hipStream_t stream)
{
    return rhs_impl(plan, out, in, stream, nullptr);
}

int sgr_rhs_events(sgr_plan *plan, double *const out[SGR_NUM_VARS], const double *const in[SGR_NUM_VARS],
                   hipStream_t stream, hipEvent_t events[4])
{
    if (!events) return fail(SGR_E_INVALID, "events is null");
    return rhs_impl(plan, out, in, stream, events);
}

int sgr_constraints(sgr_plan *plan, const double *const in[SGR_NUM_VARS], double *const out[4], hipStream_t stream)
{
    if (!plan || !in || !out) return fail(SGR_E_INVALID, "null argument");
    if (plan->variant == Variant::Fused)
        return fail(SGR_E_UNSUPPORTED, "constraints need the two-pass kernels (unset SGR_KERNEL=fused)");
    LaunchArgs a{};
    OutPtrs co{};
    for (int f = 0; f < NF; f++) {
        if (!in[f]) return fail(SGR_E_INVALID, "null field pointer");
        a.in.p[f] = in[f];
    }
    for (int c = 0; c < 4; c++) {
        if (!out[c]) return fail(SGR_E_INVALID, "null constraint pointer");
        co.p[c] = out[c];
    }
    a.blocks = plan->d_blocks;
    a.tile_block = plan->d_tile_block;
    a.ntiles = plan->ntiles;
    a.ws = plan->d_ws;
    a.wsK = plan->d_wsK;
    a.np = plan->np;
    a.upwind = plan->params.adv_mode == SGR_ADV_UPWIND;
    a.mixed = plan->params.precision == SGR_PREC_MIXED;
    a.derivs_threads = plan->derivs_threads;
    int cur = 0;
    hipGetDevice(&cur);
    if (cur != plan->device) HIP_TRY(hipSetDevice(plan->device), "hipSetDevice");
    hipError_t e = hipSuccess;
    for (uint32_t t0 = 0; e == hipSuccess && t0 < plan->ntiles; t0 += plan->chunk_tiles) {
        a.tile0 = t0;
        a.ntiles = plan->ntiles - t0 < plan->chunk_tiles ? plan->ntiles - t0 : plan->chunk_tiles;
        e = launch_constraints(a, co, stream);
    }
    if (cur != plan->device) hipSetDevice(cur);
    if (e != hipSuccess) return hip_fail(e, "constraint kernels");
    return SGR_OK;
}

int sgr_ghost_fill(sgr_plan *plan, double *const u[SGR_NUM_VARS], hipStream_t stream)
{
    if (!plan || !u) return fail(SGR_E_INVALID, "null argument");
    if (!plan->ghost.complete) return fail(SGR_E_UNSUPPORTED, plan->ghost.msg);
    OutPtrs o;
    for (int f = 0; f < NF; f++) {
        if (!u[f]) return fail(SGR_E_INVALID, "null field pointer");
        o.p[f] = u[f];
    }
    int cur = 0;
    hipGetDevice(&cur);
    if (cur != plan->device) HIP_TRY(hipSetDevice(plan->device), "hipSetDevice");
    hipError_t e = launch_ghost_fill(o, plan->d_ghost, plan->d_gbox, plan->ghost.bstart, stream);
    if (cur != plan->device) hipSetDevice(cur);
    if (e != hipSuccess) return hip_fail(e, "ghost fill launch");
    return SGR_OK;
}

int sgr_ghost_table_stats(const sgr_block *blocks, uint32_t nblocks, uint64_t stats[3])
{
    if (!stats || (nblocks && !blocks)) return fail(SGR_E_INVALID, "null argument");
    for (uint32_t b = 0; b < nblocks; b++) {
        const int rc = validate_block(blocks[b], b);
        if (rc) return rc;
    }
    const GhostTable T = build_ghost_table(blocks, nblocks);
    stats[0] = T.patches.size();
    stats[1] = T.nodes;
    stats[2] = T.interp_nodes;
    if (!T.complete) return fail(SGR_E_UNSUPPORTED, T.msg);
    return SGR_OK;
}

int sgr_hbm_copy(void *dst, const void *src, uint64_t bytes, hipStream_t stream)
{
    if (!dst || !src) return fail(SGR_E_INVALID, "null argument");
    if (bytes % 16 || (uintptr_t)dst % 16 || (uintptr_t)src % 16)
        return fail(SGR_E_INVALID, "pointers and size must be multiples of 16 bytes");
    if (bytes == 0) return SGR_OK;
    const uint64_t n = bytes / 16;
    const uint32_t grid = (uint32_t)std::min<uint64_t>((n + 255) / 256, 256ull * 64);
    hipLaunchKernelGGL(k_hbm_copy, dim3(grid), dim3(256), 0, stream, (double2 *)dst, (const double2 *)src, n);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "copy launch");
    return SGR_OK;
}

int sgr_rk4_stage(sgr_plan *plan, int stage, double dt, const double *const u_n[SGR_NUM_VARS],
                  const double *const u_stage[SGR_NUM_VARS], double *const acc[SGR_NUM_VARS],
                  double *const u_next[SGR_NUM_VARS], hipStream_t stream)
{
    if (!plan || !u_n || !u_stage || !acc || !u_next) return fail(SGR_E_INVALID, "null argument");
    if (stage < 0 || stage > 3) return fail(SGR_E_INVALID, "stage must be 0..3");
    if (!std::isfinite(dt)) return fail(SGR_E_INVALID, "dt must be finite");
    RkArgs rk{};
    rk.stage = stage;
    const double cn[3] = {dt / 2, dt / 2, dt};
    const double w[4] = {dt / 6, dt / 3, dt / 3, dt / 6};
    rk.cn = stage < 3 ? cn[stage] : 0.0;
    rk.w = w[stage];
    for (int f = 0; f < NF; f++) {
        if (!u_n[f] || !acc[f] || !u_next[f]) return fail(SGR_E_INVALID, "null field pointer");
        rk.u0.p[f] = u_n[f];
        rk.acc.p[f] = acc[f];
        rk.next.p[f] = u_next[f];
    }
    return rhs_impl(plan, u_next, u_stage, stream, nullptr, &rk);
}

int sgr_rk4_step(sgr_plan *plan, double dt, double *const u[SGR_NUM_VARS], double *const s_a[SGR_NUM_VARS],
                 double *const s_b[SGR_NUM_VARS], double *const acc[SGR_NUM_VARS], hipStream_t stream)
{
    if (!plan || !u || !s_a || !s_b || !acc) return fail(SGR_E_INVALID, "null argument");
    int rc;
    if ((rc = sgr_ghost_fill(plan, u, stream))) return rc;
    if ((rc = sgr_rk4_stage(plan, 0, dt, u, u, acc, s_a, stream))) return rc;
    if ((rc = sgr_ghost_fill(plan, s_a, stream))) return rc;
    if ((rc = sgr_rk4_stage(plan, 1, dt, u, s_a, acc, s_b, stream))) return rc;
    if ((rc = sgr_ghost_fill(plan, s_b, stream))) return rc;
    if ((rc = sgr_rk4_stage(plan, 2, dt, u, s_b, acc, s_a, stream))) return rc;
    if ((rc = sgr_ghost_fill(plan, s_a, stream))) return rc;
    return sgr_rk4_stage(plan, 3, dt, u, s_a, acc, u, stream);
}

struct sgr_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int device = 0;
};

int sgr_rk4_graph_create(sgr_plan *plan, double dt, double *const u[SGR_NUM_VARS], double *const s_a[SGR_NUM_VARS],
                         double *const s_b[SGR_NUM_VARS], double *const acc[SGR_NUM_VARS], sgr_graph **graph)
{
    if (!plan || !graph) return fail(SGR_E_INVALID, "null argument");
    *graph = nullptr;
    int cur = 0;
    hipGetDevice(&cur);
    HIP_TRY(hipSetDevice(plan->device), "hipSetDevice");
    hipStream_t s = nullptr;
    hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e != hipSuccess) {
        hipSetDevice(cur);
        return hip_fail(e, "hipStreamCreate");
    }
    sgr_graph *g = new sgr_graph;
    g->device = plan->device;
    int rc = SGR_OK;
    e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) rc = hip_fail(e, "hipStreamBeginCapture");
    if (rc == SGR_OK) {
        rc = sgr_rk4_step(plan, dt, u, s_a, s_b, acc, s);
        e = hipStreamEndCapture(s, &g->graph);
        if (rc == SGR_OK && e != hipSuccess) rc = hip_fail(e, "hipStreamEndCapture");
    }
    if (rc == SGR_OK) {
        e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
        if (e != hipSuccess) rc = hip_fail(e, "hipGraphInstantiate");
    }
    hipStreamDestroy(s);
    hipSetDevice(cur);
    if (rc != SGR_OK) {
        if (g->exec) hipGraphExecDestroy(g->exec);
        if (g->graph) hipGraphDestroy(g->graph);
        delete g;
        return rc;
    }
    *graph = g;
    return SGR_OK;
}

int sgr_graph_launch(sgr_graph *graph, hipStream_t stream)
{
    if (!graph) return fail(SGR_E_INVALID, "null graph");
    HIP_TRY(hipGraphLaunch(graph->exec, stream), "hipGraphLaunch");
    return SGR_OK;
}

int sgr_graph_destroy(sgr_graph *graph)
{
    if (!graph) return SGR_OK;
    if (graph->exec) hipGraphExecDestroy(graph->exec);
    if (graph->graph) hipGraphDestroy(graph->graph);
    delete graph;
    return SGR_OK;
}

int sgr_check_finite(sgr_plan *plan, const double *const in[SGR_NUM_VARS], hipStream_t stream)
{
    if (!plan || !in) return fail(SGR_E_INVALID, "null argument");
    FieldPtrs fp;
    for (int f = 0; f < NF; f++) {
        if (!in[f]) return fail(SGR_E_INVALID, "null field pointer");
        fp.p[f] = in[f];
    }
    if (plan->hblocks.empty()) return SGR_OK;
    int cur = 0;
    hipGetDevice(&cur);
    HIP_TRY(hipSetDevice(plan->device), "hipSetDevice");
    int h = 0;
    hipError_t e = hipMemsetAsync(plan->d_flag, 0, sizeof(int), stream);
    if (e == hipSuccess) {
        const uint32_t nb = (uint32_t)plan->hblocks.size();
        hipLaunchKernelGGL(k_check_finite, dim3(64, nb < 65535 ? nb : 65535), dim3(256), 0, stream, fp, plan->d_blocks,
                           nb, plan->d_flag);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(&h, plan->d_flag, sizeof(int), hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    hipSetDevice(cur);
    if (e != hipSuccess) return hip_fail(e, "check_finite");
    if (h) return fail(SGR_E_NONFINITE, "non-finite value in the padded input of at least one block");
    return SGR_OK;
}

}  // extern "C"

// Per-block entry points (sgr_rhs_block, sgr_bssnrhs_host).  A Dendro driver
// calls bssnrhs block after block, so the plan and the staging buffers are
// kept between calls in one process-wide context, as the reference keeps its
// one memory pool: a one-block plan, rebuilt only when the device, the block
// size or the kernel-selecting parameters (advection, precision, SGR_KERNEL)
// change; the block's offset, coordinates and bflag are re-uploaded per call.
// Calls are serialised by a mutex.  The buffers are held until the process
// ends.
namespace {
struct BlockStage {
    std::mutex m;
    sgr_plan *plan = nullptr;
    int device = -1;
    uint32_t sz[3] = {0, 0, 0};
    int32_t adv = -1, prec = -1;
    uint64_t cap = 0;           // points per field held by d_in, d_out, h_out
    double *d_in = nullptr, *d_out = nullptr;
    double *h_out = nullptr;    // pinned
};
BlockStage g_stage;

// The plan of the staging context set up for block b (sz, offset, geometry)
// and params on `device`; the caller holds g_stage.m and the device is current.
int stage_plan(const sgr_block &b, const sgr_params *params, int device)
{
    BlockStage &S = g_stage;
    const char *kv = getenv("SGR_KERNEL");
    const Variant want = kv && !strcmp(kv, "fused") ? Variant::Fused : Variant::Workspace;
    const bool same = S.plan && S.device == device && S.sz[0] == b.sz[0] && S.sz[1] == b.sz[1] &&
                      S.sz[2] == b.sz[2] && S.adv == params->adv_mode && S.prec == params->precision &&
                      S.plan->variant == want;
    if (!same) {
        if (S.plan) sgr_plan_destroy(S.plan);
        S.plan = nullptr;
        const int rc = sgr_plan_create(&S.plan, &b, 1, params, device);
        if (rc) return rc;
        S.device = device;
        for (int d = 0; d < 3; d++) S.sz[d] = b.sz[d];
        S.adv = params->adv_mode;
        S.prec = params->precision;
        return SGR_OK;
    }
    // same size: new offset, coordinates, bflag and physics parameters
    if (S.plan->variant == Variant::Fused && (params->gauge != SGR_GAUGE_STANDARD || params->eta_mode != SGR_ETA_CONST))
        return fail(SGR_E_UNSUPPORTED, "the fused kernel evaluates the standard gauge with constant eta only");
    BlockDesc &d = S.plan->hblocks[0];
    d.offset = b.offset;
    d.bflag = b.bflag;
    for (int k = 0; k < 3; k++) {
        d.pmin[k] = b.pmin[k];
        d.ax[k] = axis_const(b.pmin[k], b.pmax[k], b.sz[k]);
    }
    S.plan->params = *params;
    HIP_TRY(hipMemcpy(S.plan->d_blocks, &d, sizeof(BlockDesc), hipMemcpyHostToDevice), "hipMemcpy(block)");
    return SGR_OK;
}

int stage_buffers(uint64_t n, int device)
{
    BlockStage &S = g_stage;
    if (S.d_in && S.cap >= n && S.device == device) return SGR_OK;
    hipFree(S.d_in);
    hipFree(S.d_out);
    hipHostFree(S.h_out);
    S.d_in = S.d_out = S.h_out = nullptr;
    S.cap = 0;
    hipError_t e = hipMalloc(&S.d_in, n * NF * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&S.d_out, n * NF * sizeof(double));
    if (e == hipSuccess) e = hipHostMalloc(&S.h_out, n * NF * sizeof(double), hipHostMallocDefault);
    if (e != hipSuccess) return hip_fail(e, "staging buffers");
    S.cap = n;
    return SGR_OK;
}

int block_checks(const double pmin[3], const double pmax[3], const uint32_t sz[3], uint32_t bflag,
                 const sgr_params *params, uint64_t offset, sgr_block &b)
{
    if (!pmin || !pmax || !sz || !params) return fail(SGR_E_INVALID, "null argument");
    int rc = validate_params(params);
    if (rc) return rc;
    b.offset = offset;
    b.bflag = bflag;
    for (int d = 0; d < 3; d++) {
        b.sz[d] = sz[d];
        b.pmin[d] = pmin[d];
        b.pmax[d] = pmax[d];
    }
    return validate_block(b, 0);
}
}  // namespace

extern "C" {

int sgr_rhs_block(double *const out[SGR_NUM_VARS], const double *const in[SGR_NUM_VARS], uint64_t offset,
                  const double pmin[3], const double pmax[3], const uint32_t sz[3], uint32_t bflag,
                  const sgr_params *params, int device)
{
    if (!out || !in) return fail(SGR_E_INVALID, "null argument");
    sgr_block b;
    int rc = block_checks(pmin, pmax, sz, bflag, params, offset, b);
    if (rc) return rc;
    std::lock_guard<std::mutex> lock(g_stage.m);
    int cur = 0;
    hipGetDevice(&cur);
    HIP_TRY(hipSetDevice(device), "hipSetDevice");
    rc = stage_plan(b, params, device);
    if (rc == SGR_OK) rc = sgr_rhs(g_stage.plan, out, in, nullptr);
    if (rc == SGR_OK) {
        hipError_t e = hipStreamSynchronize(nullptr);
        if (e != hipSuccess) rc = hip_fail(e, "hipStreamSynchronize");
    }
    hipSetDevice(cur);
    return rc;
}

int sgr_bssnrhs_host(double *const out[SGR_NUM_VARS], const double *const in[SGR_NUM_VARS], uint64_t offset,
                     const double pmin[3], const double pmax[3], const uint32_t sz[3], uint32_t bflag,
                     const sgr_params *params, int device)
{
    if (!out || !in) return fail(SGR_E_INVALID, "null argument");
    for (int f = 0; f < NF; f++)
        if (!in[f] || !out[f]) return fail(SGR_E_INVALID, "null field pointer");
    sgr_block b;
    int rc = block_checks(pmin, pmax, sz, bflag, params, 0, b);  // staged at offset 0
    if (rc) return rc;
    const uint64_t n = (uint64_t)sz[0] * sz[1] * sz[2];
    std::lock_guard<std::mutex> lock(g_stage.m);
    int cur = 0;
    hipGetDevice(&cur);
    HIP_TRY(hipSetDevice(device), "hipSetDevice");
    rc = stage_buffers(n, device);
    if (rc == SGR_OK) rc = stage_plan(b, params, device);
    hipError_t e = hipSuccess;
    for (int f = 0; f < NF && rc == SGR_OK && e == hipSuccess; f++)
        e = hipMemcpy(g_stage.d_in + f * n, in[f] + offset, n * sizeof(double), hipMemcpyHostToDevice);
    if (rc == SGR_OK && e != hipSuccess) rc = hip_fail(e, "staging input");
    if (rc == SGR_OK) {
        double *o[NF];
        const double *i[NF];
        for (int f = 0; f < NF; f++) {
            o[f] = g_stage.d_out + f * n;
            i[f] = g_stage.d_in + f * n;
        }
        rc = sgr_rhs(g_stage.plan, o, i, nullptr);
    }
    if (rc == SGR_OK) {
        e = hipMemcpy(g_stage.h_out, g_stage.d_out, n * NF * sizeof(double), hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = hip_fail(e, "copying RHS back");
    }
    hipSetDevice(cur);
    if (rc) return rc;
    // only interior points are written, as the reference does
    const uint32_t nx = sz[0], ny = sz[1], nz = sz[2];
    for (int f = 0; f < NF; f++)
        for (uint32_t k = 3; k < nz - 3; k++)
            for (uint32_t j = 3; j < ny - 3; j++) {
                const uint64_t row = (uint64_t)nx * (j + (uint64_t)ny * k);
                std::memcpy(out[f] + offset + row + 3, g_stage.h_out + f * n + row + 3, (nx - 6) * sizeof(double));
            }
    return SGR_OK;
}

size_t sgr_last_error(char *buf, size_t len)
{
    if (buf && len) {
        const size_t m = g_err.size() < len - 1 ? g_err.size() : len - 1;
        std::memcpy(buf, g_err.data(), m);
        buf[m] = 0;
    }
    return g_err.size();
}

const char *sgr_version(void) { return "sympygr_amd 0.1 (gfx950)"; }

}  // extern "C"
