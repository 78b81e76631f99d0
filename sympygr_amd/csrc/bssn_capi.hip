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

// The pipelined schedule: up to PIPE_CHUNKS chunks of at least
// PIPE_MIN_TILES tiles, or as many as the workspace budget needs.  Measured:
// the 3032-leaf octree (233,738 tiles) ran 10-12% faster with 16 pipelined
// chunks than with its 3 consecutive ones and 4-8% faster than in one chunk;
// the 216 interior-only 32^3 blocks (13,824 tiles) 2-3% faster in 4 chunks
// than in one, but slower in 8 (mixed precision); chunks of 2048 tiles
// (512 x 32^3 in 16) lost up to 18%.
constexpr uint64_t PIPE_CHUNKS = 16;
constexpr uint64_t PIPE_MIN_TILES = 4096;

// Workspace bytes per 8^3 tile: RHS pass (staged derivatives in W plus the
// double tails) and constraint pass (its derivative slots plus KO sums).
uint64_t rhs_tile_bytes(size_t wsz, int nf)
{
    return (uint64_t)TILE_PTS * (rs_nslot(nf) * wsz + nf * sizeof(double));
}
uint64_t con_tile_bytes(size_t wsz) { return (uint64_t)TILE_PTS * (NSLOT * wsz + NF * sizeof(double)); }

// RHS tile shapes tuned per interior size (gen/tile_table.inc, written by
// `python -m sympygr_amd.autotune`); a size not in the table uses 8^3.
struct TileShapeEntry {
    uint16_t n[3];  // interior points per axis
    uint8_t t[3];   // tile extents
};
const TileShapeEntry kTileTable[] = {
#include "gen/tile_table.inc"
    {{0, 0, 0}, {T, T, T}},
};

// The RHS tile shape of a block: SGR_TILE_SHAPE=tx,ty,tz (every block; for
// tuning and tests), else the table, else 8^3.
int tile_shape(const sgr_block &b, int td[3])
{
    td[0] = td[1] = td[2] = T;
    if (const char *sv = getenv("SGR_TILE_SHAPE")) {
        if (sscanf(sv, "%d,%d,%d", &td[0], &td[1], &td[2]) != 3 || td[0] < 2 || td[1] < 2 || td[2] < 2 ||
            !tile_shape_ok(td[0], td[1], td[2]))
            return fail(SGR_E_INVALID, std::string("SGR_TILE_SHAPE=") + sv +
                                           " is not a tile shape (tx,ty,tz >= 2, at most 512 points, within the "
                                           "stencil kernel's LDS arrays)");
        return SGR_OK;
    }
    for (const TileShapeEntry &e : kTileTable) {
        if (e.n[0] == b.sz[0] - 6 && e.n[1] == b.sz[1] - 6 && e.n[2] == b.sz[2] - 6) {
            for (int k = 0; k < 3; k++) td[k] = e.t[k];
            break;
        }
    }
    return SGR_OK;
}

// The kernels' descriptor of one validated block with RHS tile extents td,
// whose first RHS tile has the global id tile_base and first constraint-pass
// tile the id tile_base8.
BlockDesc block_desc(const sgr_block &b, const int td[3], uint32_t tile_base, uint32_t tile_base8)
{
    BlockDesc d{};
    d.offset = b.offset;
    d.bflag = b.bflag;
    d.tile_base = tile_base;
    d.tile_base8 = tile_base8;
    for (int k = 0; k < 3; k++) {
        d.n[k] = (int32_t)b.sz[k];
        d.tdim[k] = td[k];
        d.ntile[k] = (int32_t)((b.sz[k] - 6 + td[k] - 1) / td[k]);
        d.ntile8[k] = (int32_t)((b.sz[k] - 6 + T - 1) / T);
        d.pmin[k] = b.pmin[k];
        d.ax[k] = axis_const(b.pmin[k], b.pmax[k], b.sz[k]);
    }
    return d;
}
bool is_cube8(const BlockDesc &d) { return d.tdim[0] == T && d.tdim[1] == T && d.tdim[2] == T; }

// RHS tiles of a block that reach a flagged face (they take the per-point
// closure forms and cost more than the others).
uint64_t face_tiles(const BlockDesc &d)
{
    uint64_t n = 0;
    for (int tz = 0; tz < d.ntile[2]; tz++)
        for (int ty = 0; ty < d.ntile[1]; ty++)
            for (int tx = 0; tx < d.ntile[0]; tx++) {
                const int t[3] = {tx, ty, tz};
                bool f = false;
                for (int k = 0; k < 3; k++)
                    f = f || ((d.bflag >> (2 * k) & 1u) && t[k] == 0) ||
                        ((d.bflag >> (2 * k + 1) & 1u) && t[k] == d.ntile[k] - 1);
                n += f;
            }
    return n;
}

// Block order of one kind's RHS tiles in a pipelined plan: blocks grouped by
// their count of face tiles and the groups interleaved by smooth weighted
// round robin (a group of w blocks out of W is picked w times in W, evenly
// spread), so that every workspace chunk holds about the same share of the
// costlier face tiles and the two streams' passes stay in step (the
// 3032-leaf octree: 1-3% faster than block order).  Plans of one chunk keep
// block order, which keeps the blocks a launch works on at once close in
// memory (512 blocks of 32^3: 1% slower spread).  SGR_TILE_ORDER=block|spread
// decides instead.
std::vector<uint32_t> spread_order(const std::vector<uint32_t> &kind, const std::vector<BlockDesc> &hb)
{
    std::map<uint64_t, std::vector<uint32_t>> groups;
    for (uint32_t b : kind) groups[face_tiles(hb[b])].push_back(b);
    if (groups.size() <= 1) return kind;
    struct Group {
        const std::vector<uint32_t> *v;
        size_t next;
        int64_t cur;
    };
    std::vector<Group> g;
    for (const auto &kv : groups) g.push_back({&kv.second, 0, 0});
    const int64_t W = (int64_t)kind.size();
    std::vector<uint32_t> out;
    out.reserve(kind.size());
    for (int64_t i = 0; i < W; i++) {
        Group *best = nullptr;
        for (Group &x : g) {
            x.cur += (int64_t)x.v->size();
            if (x.next < x.v->size() && (!best || x.cur > best->cur)) best = &x;
        }
        best->cur -= W;
        out.push_back((*best->v)[best->next++]);
    }
    return out;
}

// Whether the stencil pass runs as a persistent grid fed by tile queues:
// when at least a tenth of the tiles reach a flagged face, the queues
// balance their extra cost (measured: 512 blocks of 32^3, 18% face tiles,
// 2-4% faster; interior-only blocks and the octree, 0-2%, 2-4% slower with
// the queues).  SGR_STENCIL_QUEUE=0|1 decides instead.
bool stencil_queue(uint64_t faces, uint64_t tiles)
{
    if (const char *v = getenv("SGR_STENCIL_QUEUE")) return atoi(v) != 0;
    return tiles && faces * 10 >= tiles;
}

// Stencil workgroups resident at once on `device`: two per CU (LDS and
// registers of the 512-thread kernel), rounded down to a multiple of the 8
// XCDs' queues; 512 (MI355X: 256 CUs) if the attribute is not available.
uint32_t resident_stencil_groups(int device)
{
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus <= 0)
        return 512;
    const uint32_t g = (uint32_t)(2 * cus) & ~7u;
    return g >= 8 ? g : 8;
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
    if (p->formulation != SGR_FORM_BSSN && p->formulation != SGR_FORM_CCZ4)
        return fail(SGR_E_INVALID, "formulation must be SGR_FORM_BSSN or SGR_FORM_CCZ4");
    if (p->formulation == SGR_FORM_CCZ4 && p->gauge != SGR_GAUGE_STANDARD)
        return fail(SGR_E_UNSUPPORTED, "CCZ4 is evolved with the standard gauge only");
    if (p->dissipation != SGR_DISS_KO && p->dissipation != SGR_DISS_TVB3 && p->dissipation != SGR_DISS_TVB5)
        return fail(SGR_E_INVALID, "dissipation must be SGR_DISS_KO, SGR_DISS_TVB3 or SGR_DISS_TVB5");
    if (p->formulation == SGR_FORM_CCZ4 && p->dissipation != SGR_DISS_KO)
        return fail(SGR_E_UNSUPPORTED, "TVB dissipation is defined for the BSSN system only");
    for (int i = 0; i < 3; i++)
        if (!std::isfinite(p->kappa[i])) return fail(SGR_E_INVALID, "kappa must be finite");
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
    for (int i = 0; i < 3; i++) q.kappa[i] = p.kappa[i];
    return q;
}

__global__ void k_check_finite(FieldPtrs in, int nf, const BlockDesc *blocks, uint32_t nblocks, int *flag)
{
    for (uint32_t b = blockIdx.y; b < nblocks; b += gridDim.y) {
        const BlockDesc &B = blocks[b];
        const uint64_t n = (uint64_t)B.n[0] * B.n[1] * B.n[2];
        for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < n; e += (uint64_t)gridDim.x * blockDim.x) {
            bool bad = false;
            for (int f = 0; f < nf; f++) bad |= !isfinite(in.p[f][B.offset + e]);
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

struct sgr_plan {
    int derivs_threads = 512;  // constraint pass: k_derivs workgroup size (512: one point per thread, 256: two)
    int device = 0;
    sgr_params params{};
    std::vector<BlockDesc> hblocks;
    BlockDesc *d_blocks = nullptr;
    uint32_t *d_tile_block = nullptr;   // RHS tiles: the 8^3-shaped blocks' tiles first
    uint32_t *d_tile_block8 = nullptr;  // constraint-pass tiles (8^3), block order
    uint32_t ntiles = 0;
    uint32_t nfast = 0;                 // RHS tiles of 8^3-shaped blocks (tile ids 0 .. nfast-1)
    uint32_t ntiles8 = 0;
    uint32_t *d_queue = nullptr;        // 8 tile-queue counters of the stencil pass
    bool use_queue = false;             // stencil pass as a persistent grid fed by the tile queues
    uint32_t grid_cap = 512;            // resident stencil workgroups: 2 per CU, a multiple of 8
    // One device workspace, tile-major, shared by the two passes that use it:
    //   RHS:         chunk_tiles tiles of rs_nslot(nf) W values per point,
    //                then their nf double tails per point
    //   constraints: cchunk_tiles tiles of NSLOT W values per point, then
    //                their NF double KO sums per point
    void *d_ws = nullptr;
    int *d_flag = nullptr;
    // TVB dissipation: characteristic speeds along x, y, z at every node of
    // the field arrays (lam_span nodes each, the end of the last block)
    double *d_lam = nullptr;
    uint64_t lam_span = 0;
    // Pipelined schedule (two workspace buffers): the stencil pass of chunk
    // c + 1 on the caller's stream overlaps the pointwise pass of chunk c on
    // `side`; pev: fork, join, stencil done (per buffer), pointwise done (per buffer)
    bool pipe = false;
    hipStream_t side = nullptr;
    hipEvent_t pev[6] = {};
    uint32_t chunk_tiles = 0;    // tiles per RHS launch (workspace capacity)
    uint32_t cchunk_tiles = 0;   // tiles per constraint launch
    uint64_t ws_bytes = 0;
    uint64_t interior = 0;
    size_t wsz = sizeof(double); // bytes of one staged derivative (W)
    int nf = NF;                 // fields: NF (BSSN) or NF_CCZ4
    // ghost refresh table
    GhostTable ghost;
    GhostPatch *d_ghost = nullptr;
    GhostBox *d_gbox = nullptr;
};

static void plan_free(sgr_plan *p)
{
    if (!p) return;
    int cur = 0;
    hipGetDevice(&cur);
    hipSetDevice(p->device);
    hipFree(p->d_blocks);
    hipFree(p->d_tile_block);
    hipFree(p->d_tile_block8);
    hipFree(p->d_queue);
    hipFree(p->d_ws);
    hipFree(p->d_flag);
    hipFree(p->d_lam);
    if (p->side) {
        hipStreamSynchronize(p->side);
        hipStreamDestroy(p->side);
    }
    for (hipEvent_t &e : p->pev)
        if (e) hipEventDestroy(e);
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
    p->formulation = SGR_FORM_BSSN;
    p->kappa[0] = 0.02;
    p->kappa[1] = 0.0;
    p->kappa[2] = 0.5;
    p->dissipation = SGR_DISS_KO;
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
    p->nf = params->formulation == SGR_FORM_CCZ4 ? NF_CCZ4 : NF;
    // SGR_DERIVS_THREADS=256|512 sets the derivative kernel's workgroup size.
    {
        const char *tv = getenv("SGR_DERIVS_THREADS");
        if (tv && !strcmp(tv, "512")) p->derivs_threads = 512;
        else if (tv && !strcmp(tv, "256")) p->derivs_threads = 256;
    }
    // RHS tiles of the 8^3-shaped blocks first, then those of other shapes
    // (two stencil kernels, each over a contiguous tile range), each kind in
    // block order or, in a pipelined plan, spread (spread_order); constraint
    // tiles (8^3) in block order
    std::vector<uint32_t> tile_block, tile_block8;
    std::vector<int> shape(3 * (size_t)nblocks);
    for (uint32_t b = 0; b < nblocks; b++) {
        rc = validate_block(blocks[b], b);
        if (!rc) rc = tile_shape(blocks[b], &shape[3 * b]);
        if (rc) {
            delete p;
            return rc;
        }
        p->hblocks.push_back(block_desc(blocks[b], &shape[3 * b], 0, 0));
        p->interior += (uint64_t)(blocks[b].sz[0] - 6) * (blocks[b].sz[1] - 6) * (blocks[b].sz[2] - 6);
        if (params->dissipation != SGR_DISS_KO)
            p->lam_span = std::max<uint64_t>(p->lam_span, blocks[b].offset + (uint64_t)blocks[b].sz[0] *
                                                                               blocks[b].sz[1] * blocks[b].sz[2]);
    }
    uint64_t nt_all = 0, nt8_all = 0, faces = 0;  // RHS and constraint tiles, RHS tiles at a flagged face
    for (const BlockDesc &d : p->hblocks) {
        nt_all += (uint64_t)d.ntile[0] * d.ntile[1] * d.ntile[2];
        nt8_all += (uint64_t)d.ntile8[0] * d.ntile8[1] * d.ntile8[2];
        faces += face_tiles(d);
    }
    p->use_queue = stencil_queue(faces, nt_all);
    // The workspace holds at most `chunk_tiles` tiles of the RHS pass (budget
    // SGR_WS_GIB, default 40 GiB); larger plans run as several launches over
    // equal tile ranges.  The constraint pass stages more per tile and runs
    // in as many launches as that same allocation needs (at least one tile).
    const size_t wsz = params->precision == SGR_PREC_MIXED ? sizeof(float) : sizeof(double);
    p->wsz = wsz;
    const uint64_t rhs_tile = rhs_tile_bytes(wsz, p->nf), con_tile = con_tile_bytes(wsz);
    {
        double gib = 40.0;
        const char *gv = getenv("SGR_WS_GIB");
        if (gv) gib = atof(gv);
        uint64_t cap = (uint64_t)(gib * 1073741824.0) / rhs_tile;
        if (cap < 1) cap = 1;
        const uint64_t nt = nt_all;
        // A plan runs pipelined (two workspace buffers, each chunk's stencil
        // pass overlapping the previous chunk's pointwise pass on a second
        // stream) when its workspace takes more than one chunk, or when it
        // has at least 2 PIPE_MIN_TILES tiles and one workgroup per stencil
        // tile (the persistent tile-queue grid holds every CU and leaves the
        // other stream nothing to overlap).  Chunks: twice as many as the
        // budget needs, or up to PIPE_CHUNKS of at least PIPE_MIN_TILES
        // tiles.  SGR_PIPE_CHUNKS=k forces the pipelined schedule with at
        // least k >= 2 chunks; SGR_PIPE_CHUNKS=0 forces consecutive chunks.
        const uint64_t need = nt ? (nt + cap - 1) / cap : 1;
        const char *pv = getenv("SGR_PIPE_CHUNKS");
        const uint64_t pk = pv ? (uint64_t)std::max(0, atoi(pv)) : 0;
        p->pipe = nt >= 2 && (pv ? pk >= 2 : need >= 2 || (!p->use_queue && nt >= 2 * PIPE_MIN_TILES));
        // equal chunks (no short last launch)
        uint64_t nchunks = need;
        if (p->pipe)
            nchunks = std::min<uint64_t>(
                nt, std::max<uint64_t>({2 * need, pv ? pk : std::min(PIPE_CHUNKS, nt / PIPE_MIN_TILES), 2}));
        p->chunk_tiles = (uint32_t)((nt + nchunks - 1) / nchunks);
        p->ws_bytes = nt ? std::max<uint64_t>((p->pipe ? 2 : 1) * p->chunk_tiles * rhs_tile, con_tile) : 0;
        p->cchunk_tiles = (uint32_t)std::min<uint64_t>(nt8_all, p->ws_bytes / con_tile);
    }
    for (int pass = 0; pass < 2; pass++) {
        std::vector<uint32_t> kind;
        for (uint32_t b = 0; b < nblocks; b++)
            if (is_cube8(p->hblocks[b]) == (pass == 0)) kind.push_back(b);
        const char *ov = getenv("SGR_TILE_ORDER");
        const bool spread = ov ? !strcmp(ov, "spread") : p->pipe;
        for (uint32_t b : spread ? spread_order(kind, p->hblocks) : kind) {
            BlockDesc &d = p->hblocks[b];
            const uint64_t nt = (uint64_t)d.ntile[0] * d.ntile[1] * d.ntile[2];
            if (tile_block.size() + nt > 0xFFFFFFF0u) {
                delete p;
                return fail(SGR_E_INVALID, "too many tiles");
            }
            d.tile_base = (uint32_t)tile_block.size();
            d.ws_base = (uint64_t)d.tile_base * TILE_PTS;
            tile_block.insert(tile_block.end(), nt, b);
        }
        if (pass == 0) p->nfast = (uint32_t)tile_block.size();
    }
    for (uint32_t b = 0; b < nblocks; b++) {
        BlockDesc &d = p->hblocks[b];
        const uint64_t nt = (uint64_t)d.ntile8[0] * d.ntile8[1] * d.ntile8[2];
        if (tile_block8.size() + nt > 0xFFFFFFF0u) {
            delete p;
            return fail(SGR_E_INVALID, "too many tiles");
        }
        d.tile_base8 = (uint32_t)tile_block8.size();
        tile_block8.insert(tile_block8.end(), nt, b);
    }
    p->ntiles = (uint32_t)tile_block.size();
    p->ntiles8 = (uint32_t)tile_block8.size();
    // optional cap on the plan's device memory (SGR_DEVICE_MEM_LIMIT_GIB):
    // a plan that would need more is refused before anything is allocated
    {
        const char *lv = getenv("SGR_DEVICE_MEM_LIMIT_GIB");
        const uint64_t need = p->ws_bytes + nblocks * sizeof(BlockDesc) +
                              (tile_block.size() + tile_block8.size()) * sizeof(uint32_t) +
                              3 * p->lam_span * sizeof(double);
        if (lv && (double)need > atof(lv) * 1073741824.0) {
            char msg[160];
            snprintf(msg, sizeof msg, "plan needs %.3f GiB of device memory, over SGR_DEVICE_MEM_LIMIT_GIB=%s",
                     need / 1073741824.0, lv);
            delete p;
            return fail(SGR_E_NOMEM, msg);
        }
    }
    int ndev = 0;
    hipError_t derr = hipGetDeviceCount(&ndev);
    if (derr != hipSuccess || device < 0 || device >= ndev) {
        delete p;
        return derr != hipSuccess ? hip_fail(derr, "hipGetDeviceCount") : fail(SGR_E_INVALID, "device index out of range");
    }

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
    // SGR_FAULT_ALLOC=k: the k-th device allocation of this call reports
    // out-of-memory without allocating (exercises the clean-up of a
    // partially built plan)
    int alloc_no = 0;
    const char *fv = getenv("SGR_FAULT_ALLOC");
    const int fault_at = fv ? atoi(fv) : 0;
    auto dev_alloc = [&](void **ptr, size_t bytes) -> hipError_t {
        if (++alloc_no == fault_at) return hipErrorOutOfMemory;
        return hipMalloc(ptr, bytes);
    };
    if (nblocks) {
        if ((e = dev_alloc((void **)&p->d_blocks, nblocks * sizeof(BlockDesc))) != hipSuccess)
            return bail(e, "hipMalloc(blocks)");
        if ((e = hipMemcpy(p->d_blocks, p->hblocks.data(), nblocks * sizeof(BlockDesc), hipMemcpyHostToDevice)) != hipSuccess)
            return bail(e, "hipMemcpy(blocks)");
        if ((e = dev_alloc((void **)&p->d_tile_block, tile_block.size() * sizeof(uint32_t))) != hipSuccess)
            return bail(e, "hipMalloc(tiles)");
        if ((e = hipMemcpy(p->d_tile_block, tile_block.data(), tile_block.size() * sizeof(uint32_t),
                           hipMemcpyHostToDevice)) != hipSuccess)
            return bail(e, "hipMemcpy(tiles)");
        if ((e = dev_alloc((void **)&p->d_tile_block8, tile_block8.size() * sizeof(uint32_t))) != hipSuccess)
            return bail(e, "hipMalloc(constraint tiles)");
        if ((e = hipMemcpy(p->d_tile_block8, tile_block8.data(), tile_block8.size() * sizeof(uint32_t),
                           hipMemcpyHostToDevice)) != hipSuccess)
            return bail(e, "hipMemcpy(constraint tiles)");
        if ((e = dev_alloc(&p->d_ws, p->ws_bytes)) != hipSuccess) return bail(e, "hipMalloc(workspace)");
    }
    if ((e = dev_alloc((void **)&p->d_flag, sizeof(int))) != hipSuccess) return bail(e, "hipMalloc(flag)");
    if ((e = dev_alloc((void **)&p->d_queue, 8 * sizeof(uint32_t))) != hipSuccess) return bail(e, "hipMalloc(queue)");
    if (p->lam_span &&
        (e = dev_alloc((void **)&p->d_lam, 3 * p->lam_span * sizeof(double))) != hipSuccess)
        return bail(e, "hipMalloc(speeds)");
    p->grid_cap = resident_stencil_groups(device);
    if (p->pipe) {
        if ((e = hipStreamCreateWithFlags(&p->side, hipStreamNonBlocking)) != hipSuccess)
            return bail(e, "hipStreamCreate");
        for (hipEvent_t &ev : p->pev)
            if ((e = hipEventCreateWithFlags(&ev, hipEventDisableTiming)) != hipSuccess) return bail(e, "hipEventCreate");
    }
    p->ghost = build_ghost_table(blocks, nblocks);
    if (!p->ghost.patches.empty()) {
        const size_t gb = p->ghost.patches.size() * sizeof(GhostPatch);
        if ((e = dev_alloc((void **)&p->d_ghost, gb)) != hipSuccess) return bail(e, "hipMalloc(ghost table)");
        if ((e = hipMemcpy(p->d_ghost, p->ghost.patches.data(), gb, hipMemcpyHostToDevice)) != hipSuccess)
            return bail(e, "hipMemcpy(ghost table)");
        const size_t bb = p->ghost.boxes.size() * sizeof(GhostBox);
        if ((e = dev_alloc((void **)&p->d_gbox, bb)) != hipSuccess) return bail(e, "hipMalloc(ghost boxes)");
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
    info->derivative_slots = rs_nslot(plan->nf) + plan->nf;
    info->chunks = plan->chunk_tiles ? (plan->ntiles + plan->chunk_tiles - 1) / plan->chunk_tiles : 0;
    info->pipelined = plan->pipe && info->chunks > 1;
    info->tile_queue = plan->use_queue;
    return SGR_OK;
}

// The chunks of a pipelined plan: chunk c uses workspace buffer c & 1; its
// stencil pass runs on `stream` once the pointwise pass of chunk c - 2 (same
// buffer) is done, its pointwise pass on plan->side once its stencil pass is.
// Fork and join through events, so a captured graph records the same
// dependencies.  events[0..1] bracket the whole call, events[2..3] are
// recorded at its end.
static hipError_t pipelined(sgr_plan *plan, LaunchArgs a, hipStream_t stream, hipEvent_t *ev)
{
    hipEvent_t *const E = plan->pev;  // fork, join, stencil done [2], pointwise done [2]
    const uint64_t buf = plan->ws_bytes / 2;
    const uint64_t wbytes = (uint64_t)plan->chunk_tiles * TILE_PTS * rs_nslot(plan->nf) * plan->wsz;
    hipError_t e = ev ? hipEventRecord(ev[0], stream) : hipSuccess;
    for (int q = 0; q < 4; q++) a.ev[q] = nullptr;
    if (e == hipSuccess) e = hipEventRecord(E[0], stream);
    if (e == hipSuccess) e = hipStreamWaitEvent(plan->side, E[0], 0);
    uint32_t c = 0;
    for (uint32_t t0 = 0; e == hipSuccess && t0 < plan->ntiles; t0 += plan->chunk_tiles, c++) {
        const int b = c & 1;
        a.tile0 = t0;
        a.ntiles = plan->ntiles - t0 < plan->chunk_tiles ? plan->ntiles - t0 : plan->chunk_tiles;
        a.nfast = plan->nfast <= t0 ? 0 : std::min(plan->nfast - t0, a.ntiles);
        a.ws = (char *)plan->d_ws + b * buf;
        a.wsT = (double *)((char *)a.ws + wbytes);
        if (c >= 2) e = hipStreamWaitEvent(stream, E[4 + b], 0);
        a.only = 1;
        if (e == hipSuccess) e = launch_rhs(a, stream);
        if (e == hipSuccess) e = hipEventRecord(E[2 + b], stream);
        if (e == hipSuccess) e = hipStreamWaitEvent(plan->side, E[2 + b], 0);
        a.only = 2;
        if (e == hipSuccess) e = launch_rhs(a, plan->side);
        if (e == hipSuccess) e = hipEventRecord(E[4 + b], plan->side);
    }
    if (e == hipSuccess) e = hipEventRecord(E[1], plan->side);
    if (e == hipSuccess) e = hipStreamWaitEvent(stream, E[1], 0);
    for (int q = 1; q < 4 && e == hipSuccess && ev; q++) e = hipEventRecord(ev[q], stream);
    return e;
}

static int rhs_impl(sgr_plan *plan, double *const out[SGR_NUM_VARS], const double *const in[SGR_NUM_VARS],
                    hipStream_t stream, hipEvent_t *ev, const RkArgs *rk = nullptr)
{
    if (!plan || !out || !in) return fail(SGR_E_INVALID, "null argument");
    LaunchArgs a{};
    for (int f = 0; f < plan->nf; f++) {
        if (!in[f] || !out[f]) return fail(SGR_E_INVALID, "null field pointer");
        a.in.p[f] = in[f];
        a.out.p[f] = out[f];
    }
    a.blocks = plan->d_blocks;
    a.tile_block = plan->d_tile_block;
    a.ntiles = plan->ntiles;
    a.nfast = plan->nfast;
    a.queue = plan->use_queue ? plan->d_queue : nullptr;
    a.grid_cap = plan->grid_cap;
    a.ws = plan->d_ws;
    a.wsT = (double *)((char *)plan->d_ws + (uint64_t)plan->chunk_tiles * TILE_PTS * rs_nslot(plan->nf) * plan->wsz);
    a.params = device_params(plan->params);
    a.upwind = plan->params.adv_mode == SGR_ADV_UPWIND;
    a.mixed = plan->params.precision == SGR_PREC_MIXED;
    a.ccz4 = plan->nf == NF_CCZ4;
    a.derivs_threads = plan->derivs_threads;
    for (int q = 0; q < 4; q++) a.ev[q] = ev ? ev[q] : nullptr;
    a.rk.stage = -1;
    if (rk) a.rk = *rk;
    if (plan->params.dissipation != SGR_DISS_KO) {
        // TVB replaces the Kreiss-Oliger term: the stencil pass leaves it out
        // of the tails and k_tvb adds sigma (D_x + D_y + D_z) to them
        a.diss = plan->params.dissipation;
        a.diss_sigma = plan->params.ko_sigma;
        a.params.ko_sigma = 0.0;
        a.lam = plan->d_lam;
        a.lam_span = plan->lam_span;
    }
    int cur = 0;
    hipGetDevice(&cur);
    if (cur != plan->device) HIP_TRY(hipSetDevice(plan->device), "hipSetDevice");
    hipError_t e = a.diss ? launch_speeds(a, stream) : hipSuccess;  // once, over every node
    if (e == hipSuccess && plan->ntiles <= plan->chunk_tiles) {
        e = launch_rhs(a, stream);
    } else if (e == hipSuccess && plan->pipe) {
        e = pipelined(plan, a, stream, ev);
    } else if (e == hipSuccess) {
        // several workspace chunks: events[0..1] bracket the whole call,
        // events[2..3] are recorded at its end
        if (ev) e = hipEventRecord(ev[0], stream);
        for (int q = 0; q < 4; q++) a.ev[q] = nullptr;
        for (uint32_t t0 = 0; e == hipSuccess && t0 < plan->ntiles; t0 += plan->chunk_tiles) {
            a.tile0 = t0;
            a.ntiles = plan->ntiles - t0 < plan->chunk_tiles ? plan->ntiles - t0 : plan->chunk_tiles;
            a.nfast = plan->nfast <= t0 ? 0 : std::min(plan->nfast - t0, a.ntiles);
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

static int constraints_impl(sgr_plan *plan, const double *const in[SGR_NUM_VARS], double *const *out, bool psi4,
                     hipStream_t stream)
{
    if (!plan || !in || !out) return fail(SGR_E_INVALID, "null argument");
    LaunchArgs a{};
    OutPtrs co{};
    for (int f = 0; f < NF; f++) {
        if (!in[f]) return fail(SGR_E_INVALID, "null field pointer");
        a.in.p[f] = in[f];
    }
    for (int c = 0; c < (psi4 ? 6 : 4); c++) {
        if (!out[c]) return fail(SGR_E_INVALID, "null constraint pointer");
        co.p[c] = out[c];
    }
    a.blocks = plan->d_blocks;
    a.tile_block = plan->d_tile_block8;
    a.ntiles = plan->ntiles8;
    a.ws = plan->d_ws;
    a.wsK = (double *)((char *)plan->d_ws + (uint64_t)plan->cchunk_tiles * TILE_PTS * NSLOT * plan->wsz);
    a.upwind = plan->params.adv_mode == SGR_ADV_UPWIND;
    a.mixed = plan->params.precision == SGR_PREC_MIXED;
    a.derivs_threads = plan->derivs_threads;
    int cur = 0;
    hipGetDevice(&cur);
    if (cur != plan->device) HIP_TRY(hipSetDevice(plan->device), "hipSetDevice");
    hipError_t e = hipSuccess;
    for (uint32_t t0 = 0; e == hipSuccess && t0 < plan->ntiles8; t0 += plan->cchunk_tiles) {
        a.tile0 = t0;
        a.ntiles = plan->ntiles8 - t0 < plan->cchunk_tiles ? plan->ntiles8 - t0 : plan->cchunk_tiles;
        e = launch_constraints(a, co, psi4, stream);
    }
    if (cur != plan->device) hipSetDevice(cur);
    if (e != hipSuccess) return hip_fail(e, "constraint kernels");
    return SGR_OK;
}

int sgr_constraints(sgr_plan *plan, const double *const in[SGR_NUM_VARS], double *const out[4], hipStream_t stream)
{
    return constraints_impl(plan, in, out, false, stream);
}

int sgr_constraints_psi4(sgr_plan *plan, const double *const in[SGR_NUM_VARS], double *const out[6],
                         hipStream_t stream)
{
    return constraints_impl(plan, in, out, true, stream);
}

int sgr_ghost_fill(sgr_plan *plan, double *const u[SGR_NUM_VARS], hipStream_t stream)
{
    if (!plan || !u) return fail(SGR_E_INVALID, "null argument");
    if (!plan->ghost.complete) return fail(SGR_E_UNSUPPORTED, plan->ghost.msg);
    OutPtrs o;
    for (int f = 0; f < plan->nf; f++) {
        if (!u[f]) return fail(SGR_E_INVALID, "null field pointer");
        o.p[f] = u[f];
    }
    int cur = 0;
    hipGetDevice(&cur);
    if (cur != plan->device) HIP_TRY(hipSetDevice(plan->device), "hipSetDevice");
    hipError_t e = launch_ghost_fill(o, plan->nf, plan->d_ghost, plan->d_gbox, plan->ghost.bstart, stream);
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
    for (int f = 0; f < plan->nf; f++) {
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
    for (int f = 0; f < plan->nf; f++) {
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
        hipLaunchKernelGGL(k_check_finite, dim3(64, nb < 65535 ? nb : 65535), dim3(256), 0, stream, fp, plan->nf,
                           plan->d_blocks, nb, plan->d_flag);
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
// calls bssnrhs block after block, often from several host threads at once
// (an OpenMP loop over blocks), so every calling thread keeps its own staging
// context: a one-block plan, a stream of its own and its device and pinned
// host buffers.  The plan is rebuilt only when the device, the block size or
// the kernel-selecting parameters (advection, precision) change; the block's
// offset, coordinates and bflag are uploaded with each call.  Contexts of
// different threads run concurrently on the GPU; each is held until its
// thread ends.
namespace {
struct BlockStage {
    sgr_plan *plan = nullptr;
    int device = -1;
    uint32_t sz[3] = {0, 0, 0};
    int32_t adv = -1, prec = -1;
    hipStream_t stream = nullptr;
    uint64_t cap = 0;            // points per field held by the buffers
    double *d_in = nullptr, *d_out = nullptr;
    double *h_in = nullptr, *h_out = nullptr;  // pinned
    BlockDesc *h_desc = nullptr;                // pinned copy of the block descriptor
    void release()
    {
        if (device >= 0) hipSetDevice(device);
        if (stream) hipStreamSynchronize(stream);
        if (plan) sgr_plan_destroy(plan);
        hipFree(d_in);
        hipFree(d_out);
        hipHostFree(h_in);
        hipHostFree(h_out);
        hipHostFree(h_desc);
        if (stream) hipStreamDestroy(stream);
        plan = nullptr;
        d_in = d_out = h_in = h_out = nullptr;
        h_desc = nullptr;
        stream = nullptr;
        cap = 0;
        device = -1;
    }
    ~BlockStage() { release(); }
};
thread_local BlockStage g_stage;

// The context's plan set up for block b (sz, offset, geometry) and params
// on `device`, and its stream; the device is current.
int stage_plan(const sgr_block &b, const sgr_params *params, int device)
{
    BlockStage &S = g_stage;
    if (S.device != device) S.release();
    if (!S.stream) {
        const hipError_t e = hipStreamCreateWithFlags(&S.stream, hipStreamNonBlocking);
        if (e != hipSuccess) return hip_fail(e, "hipStreamCreate");
        S.device = device;
    }
    if (!S.h_desc) {
        const hipError_t e = hipHostMalloc((void **)&S.h_desc, sizeof(BlockDesc), hipHostMallocDefault);
        if (e != hipSuccess) return hip_fail(e, "hipHostMalloc(block)");
    }
    const bool same = S.plan && S.sz[0] == b.sz[0] && S.sz[1] == b.sz[1] && S.sz[2] == b.sz[2] &&
                      S.adv == params->adv_mode && S.prec == params->precision &&
                      S.plan->params.formulation == params->formulation &&
                      S.plan->params.dissipation == params->dissipation;  // TVB plans hold the speed arrays
    if (!same) {
        if (S.plan) {
            hipStreamSynchronize(S.stream);
            sgr_plan_destroy(S.plan);
        }
        S.plan = nullptr;
        const int rc = sgr_plan_create(&S.plan, &b, 1, params, device);
        if (rc) return rc;
        for (int d = 0; d < 3; d++) S.sz[d] = b.sz[d];
        S.adv = params->adv_mode;
        S.prec = params->precision;
        return SGR_OK;
    }
    // same size: new offset, coordinates, bflag and physics parameters; the
    // descriptor goes up on the context's stream, ahead of the kernels
    HIP_TRY(hipStreamSynchronize(S.stream), "hipStreamSynchronize");  // the pinned copy is free again
    BlockDesc &d = S.plan->hblocks[0];
    d.offset = b.offset;
    d.bflag = b.bflag;
    for (int k = 0; k < 3; k++) {
        d.pmin[k] = b.pmin[k];
        d.ax[k] = axis_const(b.pmin[k], b.pmax[k], b.sz[k]);
    }
    S.plan->params = *params;
    S.plan->use_queue = stencil_queue(face_tiles(d), S.plan->ntiles);
    *S.h_desc = d;
    HIP_TRY(hipMemcpyAsync(S.plan->d_blocks, S.h_desc, sizeof(BlockDesc), hipMemcpyHostToDevice, S.stream),
            "hipMemcpyAsync(block)");
    return SGR_OK;
}

int stage_buffers(uint64_t n)
{
    BlockStage &S = g_stage;
    if (S.d_in && S.cap >= n) return SGR_OK;
    hipFree(S.d_in);
    hipFree(S.d_out);
    hipHostFree(S.h_in);
    hipHostFree(S.h_out);
    S.d_in = S.d_out = S.h_in = S.h_out = nullptr;
    S.cap = 0;
    hipError_t e = hipMalloc(&S.d_in, n * NF * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&S.d_out, n * NF * sizeof(double));
    if (e == hipSuccess) e = hipHostMalloc((void **)&S.h_in, n * NF * sizeof(double), hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc((void **)&S.h_out, n * NF * sizeof(double), hipHostMallocDefault);
    if (e != hipSuccess) return e == hipErrorOutOfMemory ? fail(SGR_E_NOMEM, "staging buffers: out of memory")
                                                         : hip_fail(e, "staging buffers");
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
    int cur = 0;
    hipGetDevice(&cur);
    HIP_TRY(hipSetDevice(device), "hipSetDevice");
    rc = stage_plan(b, params, device);
    if (rc == SGR_OK) rc = sgr_rhs(g_stage.plan, out, in, g_stage.stream);
    if (rc == SGR_OK) {
        hipError_t e = hipStreamSynchronize(g_stage.stream);
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
    if (params->formulation != SGR_FORM_BSSN)
        return fail(SGR_E_UNSUPPORTED, "the host-array entry (bssnrhs) evaluates the BSSN system only");
    const uint64_t n = (uint64_t)sz[0] * sz[1] * sz[2];
    int cur = 0;
    hipGetDevice(&cur);
    HIP_TRY(hipSetDevice(device), "hipSetDevice");
    rc = stage_plan(b, params, device);  // leaves the context's stream idle
    if (rc == SGR_OK) rc = stage_buffers(n);
    BlockStage &S = g_stage;
    hipError_t e = hipSuccess;
    if (rc == SGR_OK) {
        // gather the 24 padded input arrays into pinned memory, one H2D
        for (int f = 0; f < NF; f++) std::memcpy(S.h_in + f * n, in[f] + offset, n * sizeof(double));
        e = hipMemcpyAsync(S.d_in, S.h_in, n * NF * sizeof(double), hipMemcpyHostToDevice, S.stream);
        if (e != hipSuccess) rc = hip_fail(e, "staging input");
    }
    if (rc == SGR_OK) {
        double *o[NF];
        const double *i[NF];
        for (int f = 0; f < NF; f++) {
            o[f] = S.d_out + f * n;
            i[f] = S.d_in + f * n;
        }
        rc = sgr_rhs(S.plan, o, i, S.stream);
    }
    if (rc == SGR_OK) {
        e = hipMemcpyAsync(S.h_out, S.d_out, n * NF * sizeof(double), hipMemcpyDeviceToHost, S.stream);
        if (e == hipSuccess) e = hipStreamSynchronize(S.stream);
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
                std::memcpy(out[f] + offset + row + 3, S.h_out + f * n + row + 3, (nx - 6) * sizeof(double));
            }
    return SGR_OK;
}

}  // extern "C"

// Allocation-free single-block entry (sgr_rhs_block_noalloc).  Its scratch
// memory -- the block descriptor, the tile map and the derivative workspace
// of one block -- is a file-scope device array that every GPU the library
// runs on holds until the process ends.  In host code `scratch` names a
// host-side stand-in, never the device memory: its device address is taken
// once per device with hipGetSymbolAddress and only that address is used.
namespace {
constexpr uint64_t NOALLOC_SCRATCH_BYTES = 32ull << 20;  // per device
constexpr uint64_t NOALLOC_DESC_OFF = 0;                 // one BlockDesc
constexpr uint64_t NOALLOC_QUEUE_OFF = 2048;             // the stencil pass's 8 tile-queue counters
constexpr uint64_t NOALLOC_MAP_OFF = 4096;               // tile map (uint32 per tile)
constexpr uint64_t NOALLOC_WS_OFF = 65536;               // workspace (256-aligned)
constexpr uint64_t NOALLOC_MAX_TILES = (NOALLOC_WS_OFF - NOALLOC_MAP_OFF) / sizeof(uint32_t);
static_assert(sizeof(BlockDesc) <= NOALLOC_QUEUE_OFF, "block descriptor larger than its scratch slot");
__device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

std::mutex g_scratch_m;
std::vector<void *> g_scratch_ptr;  // device address of `scratch`, per device (null: not yet taken)

// Device address of this device's copy of `scratch` (the device is current).
int scratch_address(int device, unsigned char **ptr)
{
    std::lock_guard<std::mutex> lock(g_scratch_m);
    if ((size_t)device >= g_scratch_ptr.size()) g_scratch_ptr.resize((size_t)device + 1, nullptr);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
        g_scratch_ptr[device] = p;
    }
    *ptr = (unsigned char *)g_scratch_ptr[device];
    return SGR_OK;
}

// Writes the block descriptor and the tile map (every tile in block 0) into
// the scratch memory; the descriptor travels as a kernel argument.
__global__ void k_setup_block(BlockDesc d, BlockDesc *dst, uint32_t *tile_block, uint32_t ntiles)
{
    for (uint32_t t = threadIdx.x; t < ntiles; t += blockDim.x) tile_block[t] = 0;
    if (threadIdx.x == 0) *dst = d;
}
}  // namespace

extern "C" {

int sgr_rhs_block_noalloc(double *const out[SGR_NUM_VARS], const double *const in[SGR_NUM_VARS], uint64_t offset,
                          const double pmin[3], const double pmax[3], const uint32_t sz[3], uint32_t bflag,
                          const sgr_params *params, hipStream_t stream)
{
    if (!out || !in) return fail(SGR_E_INVALID, "null argument");
    sgr_block b;
    int rc = block_checks(pmin, pmax, sz, bflag, params, offset, b);
    if (rc) return rc;
    if (params->dissipation != SGR_DISS_KO)
        return fail(SGR_E_UNSUPPORTED, "TVB dissipation needs the speed arrays of a plan (sgr_plan_create)");
    const int nf = params->formulation == SGR_FORM_CCZ4 ? NF_CCZ4 : NF;
    for (int f = 0; f < nf; f++)
        if (!in[f] || !out[f]) return fail(SGR_E_INVALID, "null field pointer");
    int td[3];
    rc = tile_shape(b, td);
    if (rc) return rc;
    const BlockDesc d = block_desc(b, td, 0, 0);
    const uint64_t ntiles = (uint64_t)d.ntile[0] * d.ntile[1] * d.ntile[2];
    const size_t wsz = params->precision == SGR_PREC_MIXED ? sizeof(float) : sizeof(double);
    const uint64_t need = NOALLOC_WS_OFF + ntiles * rhs_tile_bytes(wsz, nf);
    if (ntiles > NOALLOC_MAX_TILES || need > NOALLOC_SCRATCH_BYTES) {
        char msg[200];
        snprintf(msg, sizeof msg,
                 "block of %llu tiles needs %.1f MiB of scratch; the allocation-free entry holds %llu MiB",
                 (unsigned long long)ntiles, need / 1048576.0, (unsigned long long)(NOALLOC_SCRATCH_BYTES >> 20));
        return fail(SGR_E_INVALID, msg);
    }
    int device = 0;
    HIP_TRY(hipGetDevice(&device), "hipGetDevice");
    unsigned char *sp = nullptr;
    rc = scratch_address(device, &sp);
    if (rc) return rc;
    BlockDesc *d_desc = (BlockDesc *)(sp + NOALLOC_DESC_OFF);
    uint32_t *d_map = (uint32_t *)(sp + NOALLOC_MAP_OFF);
    hipLaunchKernelGGL(k_setup_block, dim3(1), dim3(256), 0, stream, d, d_desc, d_map, (uint32_t)ntiles);
    HIP_TRY(hipGetLastError(), "k_setup_block launch");
    LaunchArgs a{};
    for (int f = 0; f < nf; f++) {
        a.in.p[f] = in[f];
        a.out.p[f] = out[f];
    }
    a.blocks = d_desc;
    a.tile_block = d_map;
    a.ntiles = (uint32_t)ntiles;
    a.nfast = is_cube8(d) ? (uint32_t)ntiles : 0;
    a.queue = stencil_queue(face_tiles(d), ntiles) ? (uint32_t *)(sp + NOALLOC_QUEUE_OFF) : nullptr;
    a.grid_cap = resident_stencil_groups(device);
    a.ws = sp + NOALLOC_WS_OFF;
    a.wsT = (double *)(sp + NOALLOC_WS_OFF + ntiles * TILE_PTS * rs_nslot(nf) * wsz);
    a.params = device_params(*params);
    a.upwind = params->adv_mode == SGR_ADV_UPWIND;
    a.mixed = params->precision == SGR_PREC_MIXED;
    a.ccz4 = nf == NF_CCZ4;
    a.rk.stage = -1;
    const hipError_t e = launch_rhs(a, stream);
    if (e != hipSuccess) return hip_fail(e, "kernel launch");
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
