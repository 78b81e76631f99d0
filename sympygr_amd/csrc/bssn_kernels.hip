// BSSN (and CCZ4) RHS kernels for gfx950 (CDNA4).
//
// The RHS is two passes over 8x8x8 interior tiles of every block, with a
// tile-major derivative workspace between them:
//
//   k_stencil    one 512-thread workgroup per tile.  Field by field, the 14^3
//                halo tile is staged in LDS (the next field's tile in flight
//                meanwhile) and every thread evaluates its point's first,
//                advective, Kreiss-Oliger and second-derivative stencils;
//                the inner derivatives of the mixed terms are staged in LDS
//                so that D_y D_x, D_z D_x and D_z D_y are the reference's
//                compositions (bssn/scripts/bssnrhs_derivs.h:7-9).  It stores
//                the derivatives the algebra reads (double, or float in mixed
//                mode; the shift and gt second derivatives already contracted
//                with the inverse metric) and each output's tail in double:
//                c_f beta.ad f + sigma KO f, or on a flagged-face point the
//                Sommerfeld value plus sigma KO.
//
//   k_pointwise  one thread per interior point: reads its fields, staged
//                derivatives and tails, runs the generated algebra of all
//                outputs in R (double or float), adds each tail in double and
//                writes the RHS or the fused RK4 stage update.
//
// Both read the block table and a tile->block map, so one launch covers any
// mix of block sizes.  k_derivs + k_constraints are the constraint pass (the
// stencil kernel of the earlier layout, which stages every first derivative);
// k_ghost_sep / k_ghost_copy refresh ghost nodes from the owning blocks.
#include <hip/hip_ext.h>

#include "bssn_kernels.h"
#include "stencils.h"

#ifndef SGR_ALGEBRA_BARRIER
#define SGR_ALGEBRA_BARRIER __builtin_amdgcn_sched_barrier(0)
#endif

namespace sgr {

struct TileCoord {
    int lt;           // tile id within the block
    int i0, j0, k0;   // padded coordinate of the tile's first interior point
    int oi, oj, ok;   // first coordinate the tile owns along each axis
};

// Tile `lt` of a block tiled with tiles of extents td[] (nt[] tiles per axis).
// The last tile along an axis is shifted back to end at the interior's end,
// so it overlaps its neighbour.  Every tile evaluates derivatives at all its
// points (each tile has its own workspace), but an output point belongs to
// exactly one tile: the one whose unshifted range [GW + td t, GW + td (t + 1))
// holds it.  Stores skip the points a shifted tile shares with the previous
// one, so read-modify-write outputs (the RK4 accumulator) are updated once.
__device__ __forceinline__ TileCoord tile_at(const BlockDesc &B, int lt, const int32_t nt[3], int d0, int d1, int d2)
{
    TileCoord c;
    c.lt = lt;
    const int tx = lt % nt[0];
    const int ty = (lt / nt[0]) % nt[1];
    const int tz = lt / (nt[0] * nt[1]);
    auto origin = [](int tt, int n, int td) {
        const int o = GW + td * tt, last = n - GW - td;
        return o < last ? o : (last > GW ? last : GW);
    };
    c.i0 = origin(tx, B.n[0], d0);
    c.j0 = origin(ty, B.n[1], d1);
    c.k0 = origin(tz, B.n[2], d2);
    c.oi = GW + d0 * tx;
    c.oj = GW + d1 * ty;
    c.ok = GW + d2 * tz;
    return c;
}

// RHS tiling (the block's tile shape) and the constraint pass's 8^3 tiling.
__device__ __forceinline__ TileCoord tile_coord(const BlockDesc &B, uint32_t tile)
{
    return tile_at(B, (int)(tile - B.tile_base), B.ntile, B.tdim[0], B.tdim[1], B.tdim[2]);
}
__device__ __forceinline__ TileCoord tile_coord8(const BlockDesc &B, uint32_t tile)
{
    return tile_at(B, (int)(tile - B.tile_base8), B.ntile8, T, T, T);
}

// Fields whose first derivatives the pointwise algebra never reads (B, At):
// apart from their advection (stored contracted) they are only needed by the
// Sommerfeld condition, i.e. on tiles that touch a flagged face.
__host__ __device__ constexpr bool d1_only_for_faces(int f) { return (f >= 9 && f <= 11) || f >= 18; }

// Linear halo element e = ex + 14 (ey + 14 ez) -> its LDS index with rows of
// 24 doubles: ex + 24 (ey + 14 ez) = e + 10 (e / 14).
__device__ __forceinline__ int lds_halo(int e) { return e + 10 * (e / TH); }

// Division of a non-negative n < 2^16 by a run-time divisor 2 <= d <= 256
// through its multiplicative inverse m = ceil(2^32 / d): floor(n / d) =
// mulhi(n, m) exactly, since n (m d - 2^32) < n d < 2^32 / d.
struct Div {
    uint32_t m;
};
__device__ __forceinline__ Div make_div(int d) { return Div{0xFFFFFFFFu / (uint32_t)d + 1u}; }
__device__ __forceinline__ int div_q(int n, const Div &v) { return (int)__umulhi((uint32_t)n, v.m); }

// NT threads per tile: 256 (two points per thread) or 512 (one point each).
template <typename W, bool UPWIND, int NT>
__global__ __launch_bounds__(NT, NT == 256 ? 3 : 4) void k_derivs(FieldPtrs in, const BlockDesc *__restrict__ blocks,
                                                const uint32_t *__restrict__ tile_block, W *__restrict__ ws,
                                                double *__restrict__ wsK, bool all_d1, uint32_t tile0)
{
    constexpr int NHALO = TH * TH * TH;            // 2744 halo elements
    // LDS rows of the halo tile are RS doubles apart: 48 dwords, so the rows
    // a wave reads in one LDS pass (8 x-values each) fall on disjoint banks
    constexpr int RS = 24, PS = RS * TH;
    constexpr int NPF = (NHALO + NT - 1) / NT;     // prefetch registers per thread
    constexpr int PPT = TILE_PTS / NT;             // points per thread
    constexpr int ZS = NT / 64;                    // z-planes per pass over the tile
    __shared__ double U[PS * TH];        // x + RS (y + 14 z), halo coordinates
    __shared__ double DX[T * TH * TH];   // D_x u: x in tile, y and z with halo
    __shared__ double DY[T * T * TH];    // D_y u: x, y in tile, z with halo
    __shared__ double GD[3 * TILE_PTS];  // d_j (d_k beta^k) accumulated over the shift fields

    const uint32_t ct = xcd_remap(blockIdx.x, gridDim.x);  // tile within this launch
    const uint32_t tile = tile0 + ct;
    const BlockDesc &B = blocks[tile_block[tile]];
    const TileCoord tc = tile_coord8(B, tile);
    const int n0 = B.n[0], n1 = B.n[1], n2 = B.n[2];
    const bool lo0 = B.bflag & 1u, hi0 = B.bflag & 2u, lo1 = B.bflag & 4u, hi1 = B.bflag & 8u;
    const bool lo2 = B.bflag & 16u, hi2 = B.bflag & 32u;
    const AxisConst &a0 = B.ax[0], &a1 = B.ax[1], &a2 = B.ax[2];
    // does this tile contain a point of a flagged face? (uniform)
    const bool face_tile = (lo0 && tc.i0 == GW) || (hi0 && tc.i0 + T >= n0 - GW) || (lo1 && tc.j0 == GW) ||
                           (hi1 && tc.j0 + T >= n1 - GW) || (lo2 && tc.k0 == GW) || (hi2 && tc.k0 + T >= n2 - GW);
    // face closures apply to the three points next to a flagged face: only
    // tiles that reach them take the per-point closure branches (uniform)
    const bool clo0 = lo0 && tc.i0 < GW + 3, chi0 = hi0 && tc.i0 + T - 1 >= n0 - GW - 3;
    const bool clo1 = lo1 && tc.j0 < GW + 3, chi1 = hi1 && tc.j0 + T - 1 >= n1 - GW - 3;
    const bool clo2 = lo2 && tc.k0 < GW + 3, chi2 = hi2 && tc.k0 + T - 1 >= n2 - GW - 3;

    const int t = threadIdx.x;
    const int lx = t & 7, ly = (t >> 3) & 7, lz = t >> 6;  // lz in [0,ZS); points lz + ZS p
    const int pi = tc.i0 + lx, pj = tc.j0 + ly;
    bool valid[PPT];
    int pk[PPT];
    uint64_t wp[PPT];
    double bv[PPT][3];  // shift at the point
    double igp[PPT][6];  // inverse conformal metric at the point, for the contracted fields
    // this tile's workspace: its 512 points contiguous within every slot
    // (layout: ws_tile_base, slot stride WS_SS)
    W *__restrict__ wsb = ws + ws_tile_base(ct, NSLOT);
    double *__restrict__ kb = wsK + ws_tile_base(ct, NF);
    constexpr uint64_t SS = WS_SS;
#pragma unroll
    for (int p = 0; p < PPT; p++) {
        pk[p] = tc.k0 + lz + ZS * p;
        valid[p] = pi < n0 - GW && pj < n1 - GW && pk[p] < n2 - GW;
        wp[p] = (uint64_t)(lx + T * (ly + T * (lz + ZS * p)));
        const uint64_t gp = B.offset + pi + (uint64_t)n0 * (pj + (uint64_t)n1 * pk[p]);
#pragma unroll
        for (int d = 0; d < 3; d++) bv[p][d] = valid[p] ? in.p[6 + d][gp] : 0.0;
        {
            // once per point instead of once per contracted field (the
            // identity outside the interior keeps the adjugate finite there)
            double g[6];
#pragma unroll
            for (int q = 0; q < 6; q++) g[q] = valid[p] ? in.p[12 + q][gp] : (q == 0 || q == 3 || q == 5 ? 1.0 : 0.0);
            inverse_metric(g, igp[p]);
        }
    }
    // halo element -> offset in the block (or -1 outside the padded block)
    int32_t src[NPF];
#pragma unroll
    for (int q = 0; q < NPF; q++) {
        const int e = t + NT * q;
        const int ex = e % TH, ey = (e / TH) % TH, ez = e / (TH * TH);
        const int gi = tc.i0 - GW + ex, gj = tc.j0 - GW + ey, gk = tc.k0 - GW + ez;
        src[q] = (e < NHALO && gi < n0 && gj < n1 && gk < n2) ? gi + n0 * (gj + n1 * gk) : -1;
    }
    {
        const double *__restrict__ s0 = in.p[0] + B.offset;
#pragma unroll
        for (int q = 0; q < NPF; q++)
            if (t + NT * q < NHALO) U[lds_halo(t + NT * q)] = src[q] >= 0 ? s0[src[q]] : 0.0;
    }
    __syncthreads();

#pragma unroll 1
    for (int f = 0; f < NF; f++) {
        // fetch the next field's halo tile into registers; it lands in LDS
        // after this field's stencils, so the HBM latency overlaps them
        double pf[NPF];
        if (f + 1 < NF) {
            const double *__restrict__ sn = in.p[f + 1] + B.offset;
#pragma unroll
            for (int q = 0; q < NPF; q++) pf[q] = src[q] >= 0 ? sn[src[q]] : 0.0;
        }
        const bool need_d1 = all_d1 || !d1_only_for_faces(f) || face_tile;
        // fields with second derivatives (need_d1 holds for all of them):
        // the tile points' own D_x u and D_y u are cells of the DX and DY
        // stages of the mixed terms, written here; same calls, same values
        const bool stage_xy = d2q(f) >= 0;
#pragma unroll
        for (int p = 0; p < PPT; p++) {
            const double *c = &U[(lx + GW) + RS * ((ly + GW) + TH * (lz + ZS * p + GW))];
            const int cx = lx + T * ((ly + GW) + TH * (lz + ZS * p + GW)), cy = lx + T * (ly + T * (lz + ZS * p + GW));
            if (!valid[p]) {
                if (stage_xy) {  // a tile point outside the interior: the stages' own conditions
                    DX[cx] = pi < n0 - GW ? d1(c, 1, pi, n0, clo0, chi0, a0) : 0.0;
                    DY[cy] = pi < n0 - GW && pj < n1 - GW ? d1(c, RS, pj, n1, clo1, chi1, a1) : 0.0;
                }
                continue;
            }
            if (need_d1 || (!UPWIND && !all_d1)) {
                const double g0 = d1(c, 1, pi, n0, clo0, chi0, a0);
                const double g1 = d1(c, RS, pj, n1, clo1, chi1, a1);
                const double g2 = d1(c, PS, pk[p], n2, clo2, chi2, a2);
                if (stage_xy) {
                    DX[cx] = g0;
                    DY[cy] = g1;
                }
                if (need_d1) {
                    ws_put(&wsb[slot_d1(f, 0) * SS + wp[p]], (W)(g0));
                    ws_put(&wsb[slot_d1(f, 1) * SS + wp[p]], (W)(g1));
                    ws_put(&wsb[slot_d1(f, 2) * SS + wp[p]], (W)(g2));
                }
                if (!UPWIND && !all_d1) {
                    SGR_EXACT
                    ws_put(&wsb[slot_adv(f) * SS + wp[p]], (W)(((bv[p][0] * g0 + bv[p][1] * g1) + bv[p][2] * g2)));
                }
            }
            if (UPWIND && !all_d1) {
                SGR_EXACT
                const double q0 = adv(c, 1, pi, n0, clo0, chi0, bv[p][0] > 0.0 ? 1.0 : 0.0, a0);
                const double q1 = adv(c, RS, pj, n1, clo1, chi1, bv[p][1] > 0.0 ? 1.0 : 0.0, a1);
                const double q2 = adv(c, PS, pk[p], n2, clo2, chi2, bv[p][2] > 0.0 ? 1.0 : 0.0, a2);
                ws_put(&wsb[slot_adv(f) * SS + wp[p]], (W)(((bv[p][0] * q0 + bv[p][1] * q1) + bv[p][2] * q2)));
            }
            {
                SGR_EXACT
                const double kx = ko(c, 1, pi, n0, clo0, chi0, a0);
                const double ky = ko(c, RS, pj, n1, clo1, chi1, a1);
                const double kz = ko(c, PS, pk[p], n2, clo2, chi2, a2);
                ws_put(&kb[f * SS + wp[p]], kx + ky + kz);
            }
        }

        if (d2q(f) >= 0) {  // uniform over the workgroup
            // the shift and gt_ij enter only contracted: igt^{lm} d_l d_m f is
            // stored, and for the shift d_j d_k beta^k is accumulated in GD
            const bool contracted = f >= 6;
            if (!contracted) {
#pragma unroll
                for (int p = 0; p < PPT; p++) {
                    if (!valid[p]) continue;
                    const double *c = &U[(lx + GW) + RS * ((ly + GW) + TH * (lz + ZS * p + GW))];
                    ws_put(&wsb[slot_d2(f, 0) * SS + wp[p]], (W)(d2(c, 1, pi, n0, clo0, chi0, a0)));
                    ws_put(&wsb[slot_d2(f, 3) * SS + wp[p]], (W)(d2(c, RS, pj, n1, clo1, chi1, a1)));
                    ws_put(&wsb[slot_d2(f, 5) * SS + wp[p]], (W)(d2(c, PS, pk[p], n2, clo2, chi2, a2)));
                }
            }
            // inner derivatives of the mixed terms, valid wherever the block
            // has them.  D_x u only where D_y and D_z read it and the tile
            // points did not write it: the y halo rows of the tile's z planes
            // (NA), and the tile's y rows of the z halo planes (NB); the
            // y-halo x z-halo corners are never read.  D_y u on the z halo
            // planes.
            constexpr int NA = T * (2 * GW) * T, NB = T * T * (2 * GW);
            for (int e = t; e < NA + NB; e += NT) {
                int x, y, z;
                if (e < NA) {
                    const int yy = (e / T) % (2 * GW);
                    x = e % T;
                    y = yy < GW ? yy : yy + T;
                    z = GW + e / (T * 2 * GW);
                } else {
                    const int r = e - NA, zz = r / (T * T);
                    x = r % T;
                    y = GW + (r / T) % T;
                    z = zz < GW ? zz : zz + T;
                }
                const int gi = tc.i0 + x, gj = tc.j0 - GW + y, gk = tc.k0 - GW + z;
                double v = 0.0;
                if (gi < n0 - GW && gj < n1 && gk < n2)
                    v = d1(&U[(x + GW) + RS * (y + TH * z)], 1, gi, n0, clo0, chi0, a0);
                DX[x + T * (y + TH * z)] = v;
            }
            for (int e = t; e < T * T * (2 * GW); e += NT) {
                const int zz = e / (T * T);
                const int x = e % T, y = (e / T) % T, z = zz < GW ? zz : zz + T;
                const int gi = tc.i0 + x, gj = tc.j0 + y, gk = tc.k0 - GW + z;
                double v = 0.0;
                if (gi < n0 - GW && gj < n1 - GW && gk < n2)
                    v = d1(&U[(x + GW) + RS * ((y + GW) + TH * z)], RS, gj, n1, clo1, chi1, a1);
                DY[x + T * (y + T * z)] = v;
            }
            __syncthreads();
            if (!contracted) {
#pragma unroll
                for (int p = 0; p < PPT; p++) {
                    if (!valid[p]) continue;
                    const double *cx = &DX[lx + T * ((ly + GW) + TH * (lz + ZS * p + GW))];
                    const double *cy = &DY[lx + T * (ly + T * (lz + ZS * p + GW))];
                    ws_put(&wsb[slot_d2(f, 1) * SS + wp[p]], (W)(d1(cx, T, pj, n1, clo1, chi1, a1)));
                    ws_put(&wsb[slot_d2(f, 2) * SS + wp[p]], (W)(d1(cx, T * TH, pk[p], n2, clo2, chi2, a2)));
                    ws_put(&wsb[slot_d2(f, 4) * SS + wp[p]], (W)(d1(cy, T * T, pk[p], n2, clo2, chi2, a2)));
                }
            } else {
#pragma unroll 1
                for (int p = 0; p < PPT; p++) {
                    if (!valid[p]) continue;
                    const double *c = &U[(lx + GW) + RS * ((ly + GW) + TH * (lz + ZS * p + GW))];
                    const double *cx = &DX[lx + T * ((ly + GW) + TH * (lz + ZS * p + GW))];
                    const double *cy = &DY[lx + T * (ly + T * (lz + ZS * p + GW))];
                    const double *ig = igp[p];
                    {
                        SGR_EXACT
                        const double h0 = d2(c, 1, pi, n0, clo0, chi0, a0);
                        const double h1 = d1(cx, T, pj, n1, clo1, chi1, a1);
                        const double h2 = d1(cx, T * TH, pk[p], n2, clo2, chi2, a2);
                        const double h3 = d2(c, RS, pj, n1, clo1, chi1, a1);
                        const double h4 = d1(cy, T * T, pk[p], n2, clo2, chi2, a2);
                        const double h5 = d2(c, PS, pk[p], n2, clo2, chi2, a2);
                        // metric_contract(ig, h)
                        const double L = ((((ig[0] * h0 + 2.0 * (ig[1] * h1)) + 2.0 * (ig[2] * h2)) + ig[3] * h3) +
                                          2.0 * (ig[4] * h4)) + ig[5] * h5;
                        ws_put(&wsb[slot_lap(f) * SS + wp[p]], (W)(L));
                        if (f <= 8) {  // shift beta^k, k = f - 6: d_j d_k beta^k for j = 0, 1, 2
                            const double q0 = f == 6 ? h0 : (f == 7 ? h1 : h2);
                            const double q1 = f == 6 ? h1 : (f == 7 ? h3 : h4);
                            const double q2 = f == 6 ? h2 : (f == 7 ? h4 : h5);
                            double *gd = &GD[wp[p]];
                            if (f == 6) {
                                gd[0] = q0;
                                gd[TILE_PTS] = q1;
                                gd[2 * TILE_PTS] = q2;
                            } else if (f == 7) {
                                gd[0] = gd[0] + q0;
                                gd[TILE_PTS] = gd[TILE_PTS] + q1;
                                gd[2 * TILE_PTS] = gd[2 * TILE_PTS] + q2;
                            } else {
                                ws_put(&wsb[slot_gdb(0) * SS + wp[p]], (W)(gd[0] + q0));
                                ws_put(&wsb[slot_gdb(1) * SS + wp[p]], (W)(gd[TILE_PTS] + q1));
                                ws_put(&wsb[slot_gdb(2) * SS + wp[p]], (W)(gd[2 * TILE_PTS] + q2));
                            }
                        }
                    }
                }
            }
        }
        __syncthreads();  // every read of U (and DX, DY) for field f is done
        if (f + 1 < NF) {
#pragma unroll
            for (int q = 0; q < NPF; q++)
                if (t + NT * q < NHALO) U[lds_halo(t + NT * q)] = pf[q];
        }
        __syncthreads();
    }
}

// Sommerfeld outgoing-radiation value of one field at a face point
// (bssn/src/rhs.cpp:833-977), evaluated in the reference's order.
__device__ __forceinline__ double sommerfeld(double x, double y, double z, double inv_r, double gx, double gy,
                                             double gz, double fval, double falloff, double asym)
{
    SGR_EXACT
    return -inv_r * (x * gx + y * gy + z * gz + falloff * (fval - asym));
}

__device__ __forceinline__ double coord(double pmin, int i, double h)
{
    SGR_EXACT
    return pmin + i * h;
}

__device__ __forceinline__ double radius(double x, double y, double z)
{
    SGR_EXACT
    return sqrt(x * x + y * y + z * z);
}

__device__ __forceinline__ double ko_add(double v, double sigma, double k)
{
    SGR_EXACT
    return v + sigma * k;
}

// Classic RK4 stage update of one field at one point (see RkArgs).
__device__ __forceinline__ void rk_store(const RkArgs &rk, int f, uint64_t pp, double k)
{
    SGR_EXACT
    const double wk = rk.w * k;
    if (rk.stage == 3) {
        out_put(&rk.next.p[f][pp], rk.acc.p[f][pp] + wk);
        return;
    }
    const double u0 = rk.u0.p[f][pp];
    out_put(&rk.acc.p[f][pp], (rk.stage == 0 ? u0 : rk.acc.p[f][pp]) + wk);
    out_put(&rk.next.p[f][pp], u0 + rk.cn * k);
}

// eta of the Gamma-driver at one point (bssn/src/rhs.cpp:206-219): the
// constant ETA_CONST, damped by (ETA_R0 / r)^ETA_DAMPING_EXP beyond ETA_R0;
// or, under eta_func (GR/bssn.py:52), BSSN_ETA_R0 sqrt(igt^ij d_i chi d_j chi)
// / (1 - chi^p0)^p1 from the point's gt, d chi and chi.
__device__ __forceinline__ double eta_const_at(const Params &P, double x, double y, double z)
{
    double e = P.eta_const;
    const double r = radius(x, y, z);
    if (r >= P.eta_r0) e *= pow(P.eta_r0 / r, P.eta_damping_exp);
    return e;
}
__device__ __forceinline__ double eta_func_at(const Params &P, const double g[6], const double dchi[3], double chi)
{
    double ig[6], h[6];
    inverse_metric(g, ig);
    h[0] = dchi[0] * dchi[0];
    h[1] = dchi[0] * dchi[1];
    h[2] = dchi[0] * dchi[2];
    h[3] = dchi[1] * dchi[1];
    h[4] = dchi[1] * dchi[2];
    h[5] = dchi[2] * dchi[2];
    return P.eta_func_r0 * sqrt(metric_contract(ig, h)) * pow(1.0 - pow(chi, P.eta_power[0]), -P.eta_power[1]);
}

__device__ __forceinline__ double tail_add(double v, double tail)
{
    SGR_EXACT
    return v + tail;
}

// First, advective, Kreiss-Oliger and (want2) pure second derivative of one
// field along one axis at one point.  A tile that reaches a flagged face on
// this axis (lo / hi, uniform over the tile) takes the per-point closure
// forms from LDS; every other tile loads the 7 taps u[-3..3] once into
// registers and evaluates the interior forms from them (same expressions,
// same order: bit-equal to the LDS forms).
template <bool UPWIND>
__device__ __forceinline__ void axis_stencils(const double *c, int s, int pos, int n, bool lo, bool hi, double beta,
                                              const AxisConst &a, bool want2, double &g, double &q, double &k,
                                              double &h)
{
    if (lo || hi) {
        g = d1(c, s, pos, n, lo, hi, a);
        q = UPWIND ? adv(c, s, pos, n, lo, hi, beta, a) : g;
        k = ko(c, s, pos, n, lo, hi, a);
        h = want2 ? d2(c, s, pos, n, lo, hi, a) : 0.0;
    } else {
        double t[7];
#pragma unroll
        for (int o = 0; o < 7; o++) t[o] = c[(o - 3) * s];
        g = d1_int(t + 3, 1, a);
        q = UPWIND ? ad_int(t + 3, 1, beta > 0.0, a) : g;
        k = ko_int(t + 3, 1, a);
        h = want2 ? d2_int(t + 3, 1, a) : 0.0;
    }
}

// k_stencil: one 512-thread workgroup per tile, one point per thread.  GEN
// false: 8^3 tiles with conflict-free padded LDS rows and constant
// indexing; GEN true: any tile shape that tile_shape_ok admits (the block's
// tdim, read at run time; threads beyond the tile's points only help load
// and stage), unpadded LDS.
// Field by field, the 14^3 halo tile is staged in LDS (the next field's tile
// in flight from HBM meanwhile) and every thread evaluates its point's
// stencils: first, advective and Kreiss-Oliger along each axis; for alpha,
// chi, the shift and gt also the second derivatives, the mixed ones as
// compositions D_y D_x, D_z D_x, D_z D_y through the LDS-staged inner
// derivatives (bssn/scripts/bssnrhs_derivs.h:7-9).  It stages for the
// pointwise kernel what the algebra reads (the RS_* slots) and each
// output's tail (its advection and Kreiss-Oliger terms, in double); on a
// flagged-face point the tail is the whole output, the Sommerfeld value
// (bssn/src/rhs.cpp:833-977) plus sigma KO.
// NFLD: 24 (BSSN) or 25 (CCZ4: Theta last, with its first derivatives staged).
// The launch covers tiles tile0 .. tile0 + ntl - 1, whose workspace tiles
// start at wsoff.
template <typename W, bool UPWIND, int NFLD, bool GEN, bool QUEUE>
__global__ __launch_bounds__(512, 4) void k_stencil(FieldPtrs in, const BlockDesc *__restrict__ blocks,
                                                     const uint32_t *__restrict__ tile_block, W *__restrict__ ws,
                                                     double *__restrict__ wsT, Params P, uint32_t tile0,
                                                     uint32_t wsoff, uint32_t ntl, uint32_t *__restrict__ queue)
{
    constexpr int NT = 512;
    // 8^3: LDS rows of the halo tile are 24 doubles apart (48 dwords), so the
    // rows a wave reads in one LDS pass (8 x-values each) fall on disjoint banks
    constexpr int NPF = ((GEN ? GEN_U_CAP : TH * TH * TH) + NT - 1) / NT;  // prefetch registers per thread
    __shared__ double U[GEN ? GEN_U_CAP : 24 * TH * TH];  // halo tile: x + RS y + PS z
    __shared__ double DX[GEN ? GEN_DX_CAP : T * TH * TH];  // D_x u: x in tile, y and z with halo
    __shared__ double DY[GEN ? GEN_DY_CAP : T * T * TH];   // D_y u: x, y in tile, z with halo
    __shared__ double GD[3 * TILE_PTS];  // d_j (d_k beta^k) accumulated over the shift fields

    // !QUEUE: one tile per workgroup, in XCD-contiguous order (the loop runs
    // once; its own instantiation keeps the single-tile kernel's register
    // allocation).  QUEUE: a persistent grid fed by per-XCD tile queues: the launch's
    // ntl tiles are split into 8 contiguous ranges, one per XCD (workgroups
    // are dispatched to XCDs round-robin by id, so workgroup b takes from
    // queue b mod 8 first, keeping neighbouring tiles, which share halo data,
    // in one L2), and a workgroup whose range is exhausted takes from the
    // others.  queue[0..7] are zero at launch; every workgroup leaves when all
    // eight ranges are exhausted.
    __shared__ uint32_t s_ct;
    for (;;) {
        uint32_t ct;  // tile within this launch
        if constexpr (QUEUE) {
            if (threadIdx.x == 0) {
                uint32_t v = 0xFFFFFFFFu;
                for (uint32_t i = 0; i < 8; i++) {
                    const uint32_t xq = (blockIdx.x + i) & 7;
                    const uint32_t lo = (uint32_t)(((uint64_t)ntl * xq) >> 3);
                    const uint32_t hi = (uint32_t)(((uint64_t)ntl * (xq + 1)) >> 3);
                    if (lo >= hi) continue;
                    const uint32_t k = atomicAdd(&queue[xq], 1u);
                    if (lo + k < hi) {
                        v = lo + k;
                        break;
                    }
                }
                s_ct = v;
            }
            __syncthreads();
            ct = s_ct;
            __syncthreads();  // every thread has read s_ct before it is written again
            if (ct == 0xFFFFFFFFu) break;
        } else {
            ct = xcd_remap(blockIdx.x, gridDim.x);
        }
        const uint32_t tile = tile0 + ct;
        const BlockDesc &B = blocks[tile_block[tile]];
        const TileCoord tc = tile_coord(B, tile);
        // tile extents, halo extents and LDS strides (constants for 8^3)
        const int TX = GEN ? B.tdim[0] : T, TY = GEN ? B.tdim[1] : T, TZ = GEN ? B.tdim[2] : T;
        const int HX = TX + 2 * GW, HY = TY + 2 * GW;
        const int RS = GEN ? HX : 24, PS = GEN ? HX * HY : 24 * TH;
        const int DXZ = TX * HY, DYZ = TX * TY;  // z strides of the D_x and D_y stages
        const int NHALO = HX * HY * (TZ + 2 * GW), NPTS = TX * TY * TZ;
        const Div dTX = make_div(TX), dTY = make_div(TY), dHX = make_div(HX), dHY = make_div(HY);
        const int n0 = B.n[0], n1 = B.n[1], n2 = B.n[2];
        const bool lo0 = B.bflag & 1u, hi0 = B.bflag & 2u, lo1 = B.bflag & 4u, hi1 = B.bflag & 8u;
        const bool lo2 = B.bflag & 16u, hi2 = B.bflag & 32u;
        const AxisConst &a0 = B.ax[0], &a1 = B.ax[1], &a2 = B.ax[2];
        // face closures apply to the three points next to a flagged face: only
        // tiles that reach them take the per-point closure branches (uniform)
        const bool clo0 = lo0 && tc.i0 < GW + 3, chi0 = hi0 && tc.i0 + TX - 1 >= n0 - GW - 3;
        const bool clo1 = lo1 && tc.j0 < GW + 3, chi1 = hi1 && tc.j0 + TY - 1 >= n1 - GW - 3;
        const bool clo2 = lo2 && tc.k0 < GW + 3, chi2 = hi2 && tc.k0 + TZ - 1 >= n2 - GW - 3;

        // In the tile-queue loop the thread index is made opaque to the compiler
        // each pass, so what derives from it (point and halo-element indices)
        // is recomputed per tile instead of hoisted out of the loop and kept
        // live across it, which spilled (scratch 144 -> 84 B/lane; general
        // kernel 2-4% faster on Dendro's 17^3 and 33^3 blocks, 8^3 kernel even).
        int t = threadIdx.x;
        if constexpr (QUEUE) asm volatile("" : "+v"(t));
        int lx, ly, lz;
        if (GEN) {
            const int r = div_q(t, dTX);
            lx = t - r * TX;
            lz = div_q(r, dTY);
            ly = r - lz * TY;
        } else {
            lx = t & 7, ly = (t >> 3) & 7, lz = t >> 6;
        }
        const bool has_pt = !GEN || t < NPTS;  // this thread has a point of the tile
        const int pi = tc.i0 + lx, pj = tc.j0 + ly, pk = tc.k0 + lz;
        const bool valid = has_pt && pi < n0 - GW && pj < n1 - GW && pk < n2 - GW;
        const bool face = valid && ((lo0 && pi == GW) || (hi0 && pi == n0 - GW - 1) || (lo1 && pj == GW) ||
                                    (hi1 && pj == n1 - GW - 1) || (lo2 && pk == GW) || (hi2 && pk == n2 - GW - 1));
        const uint64_t wp = (uint64_t)t;  // the point's position in the tile (x fastest)
        const uint64_t gp = B.offset + pi + (uint64_t)n0 * (pj + (uint64_t)n1 * pk);
        const double sigma = P.ko_sigma;
        // advection coefficient of each equation's own field (the tails)
        const double c_alpha = (double)P.lambda[0];
        const double c_beta = P.rochester ? (double)P.xi[1] : (double)P.lambda[1];
        double bv[3];     // shift at the point
        double igp[6];    // inverse conformal metric at the point (identity outside the interior)
        {
            double g[6];
    #pragma unroll
            for (int d = 0; d < 3; d++) bv[d] = valid ? in.p[6 + d][gp] : 0.0;
    #pragma unroll
            for (int q = 0; q < 6; q++) g[q] = valid ? in.p[12 + q][gp] : (q == 0 || q == 3 || q == 5 ? 1.0 : 0.0);
            inverse_metric(g, igp);
        }
        W *__restrict__ wsb = ws + ws_tile_base(ct + wsoff, rs_nslot(NFLD));
        double *__restrict__ tb = wsT + ws_tile_base(ct + wsoff, NFLD);
        constexpr uint64_t SS = WS_SS;

        // halo element -> offset in the block (or -1 outside the padded block)
        int32_t src[NPF];
    #pragma unroll
        for (int q = 0; q < NPF; q++) {
            const int e = t + NT * q;
            int ex, ey, ez;
            if (GEN) {
                const int r = div_q(e, dHX);
                ex = e - r * HX;
                ez = div_q(r, dHY);
                ey = r - ez * HY;
            } else {
                ex = e % TH, ey = (e / TH) % TH, ez = e / (TH * TH);
            }
            const int gi = tc.i0 - GW + ex, gj = tc.j0 - GW + ey, gk = tc.k0 - GW + ez;
            src[q] = (e < NHALO && gi < n0 && gj < n1 && gk < n2) ? gi + n0 * (gj + n1 * gk) : -1;
        }
        // halo element e = x + HX (y + HY z) -> its LDS place
        auto lds_at = [](int e) { return GEN ? e : lds_halo(e); };
        {
            const double *__restrict__ s0 = in.p[0] + B.offset;
    #pragma unroll
            for (int q = 0; q < NPF; q++)
                if (t + NT * q < NHALO) U[lds_at(t + NT * q)] = src[q] >= 0 ? s0[src[q]] : 0.0;
        }
        __syncthreads();

    #pragma unroll 1
        for (int f = 0; f < NFLD; f++) {
            // fetch the next field's halo tile into registers; it lands in LDS
            // after this field's stencils, so the HBM latency overlaps them
            double pf[NPF];
            if (f + 1 < NFLD) {
                const double *__restrict__ sn = in.p[f + 1] + B.offset;
    #pragma unroll
                for (int q = 0; q < NPF; q++) pf[q] = src[q] >= 0 ? sn[src[q]] : 0.0;
            }
            const bool has_d2 = d2q(f) >= 0;  // alpha, chi, the shift, gt (uniform)
            const double *c = &U[(lx + GW) + RS * (ly + GW) + PS * (lz + GW)];
            const int cx = lx + TX * (ly + GW) + DXZ * (lz + GW), cy = lx + TX * ly + DYZ * (lz + GW);
            double h0 = 0.0, h1 = 0.0, h2 = 0.0;  // pure second derivatives xx, yy, zz
            if (!valid) {
                if (has_d2 && has_pt) {  // a tile point outside the interior: the stages' own conditions
                    DX[cx] = pi < n0 - GW ? d1(c, 1, pi, n0, clo0, chi0, a0) : 0.0;
                    DY[cy] = pi < n0 - GW && pj < n1 - GW ? d1(c, RS, pj, n1, clo1, chi1, a1) : 0.0;
                }
            } else {
                double g0, g1, g2, q0, q1, q2, k0, k1, k2;
                axis_stencils<UPWIND>(c, 1, pi, n0, clo0, chi0, bv[0], a0, has_d2, g0, q0, k0, h0);
                axis_stencils<UPWIND>(c, RS, pj, n1, clo1, chi1, bv[1], a1, has_d2, g1, q1, k1, h1);
                axis_stencils<UPWIND>(c, PS, pk, n2, clo2, chi2, bv[2], a2, has_d2, g2, q2, k2, h2);
                if (has_d2) {  // the tile points' own cells of the mixed-term stages
                    DX[cx] = g0;
                    DY[cy] = g1;
                }
                double advf, kof;
                {
                    SGR_EXACT
                    advf = (bv[0] * q0 + bv[1] * q1) + bv[2] * q2;
                    kof = (k0 + k1) + k2;
                }
                if (rs_has_d1(f)) {
                    ws_put(&wsb[rs_d1(f, 0) * SS + wp], (W)g0);
                    ws_put(&wsb[rs_d1(f, 1) * SS + wp], (W)g1);
                    ws_put(&wsb[rs_d1(f, 2) * SS + wp], (W)g2);
                }
                if (f <= 1) {  // pure second derivatives of alpha and chi
                    ws_put(&wsb[rs_d2(f, 0) * SS + wp], (W)h0);
                    ws_put(&wsb[rs_d2(f, 3) * SS + wp], (W)h1);
                    ws_put(&wsb[rs_d2(f, 5) * SS + wp], (W)h2);
                }
                if (face) {
                    // the Sommerfeld value replaces the right-hand side, then + sigma KO
                    const double x = coord(B.pmin[0], pi, a0.h), y = coord(B.pmin[1], pj, a1.h);
                    const double z = coord(B.pmin[2], pk, a2.h);
                    ws_put(&tb[rs_tail(f) * SS + wp],
                           ko_add(sommerfeld(x, y, z, 1.0 / radius(x, y, z), g0, g1, g2, c[0], bc_falloff(f), bc_asym(f)),
                                  sigma, kof));
                } else if (f >= 3 && f <= 5) {
                    // Gt^i: its tail, and the Gt part of B^i's tail (completed at B^i)
                    ws_put(&tb[rs_tail(f) * SS + wp], ko_add(advf, sigma, kof));
                    {
                        SGR_EXACT
                        ws_put(&tb[rs_tail(f + 6) * SS + wp], (1.0 - (double)P.lambda[3]) * advf);
                    }
                } else if (f >= 9 && f <= 11) {
                    double v = 0.0;  // Rochester gauge: B is not evolved
                    if (!P.rochester) {
                        SGR_EXACT
                        v = ws_get(&tb[rs_tail(f) * SS + wp]) + (double)P.lambda[2] * advf;
                    }
                    ws_put(&tb[rs_tail(f) * SS + wp], ko_add(v, sigma, kof));
                } else {
                    const double cf = f == 0 ? c_alpha : (f >= 6 && f <= 8) ? c_beta : 1.0;
                    double v;
                    {
                        SGR_EXACT
                        v = cf * advf;
                    }
                    ws_put(&tb[rs_tail(f) * SS + wp], ko_add(v, sigma, kof));
                }
            }

            if (has_d2) {  // uniform over the workgroup
                // inner derivatives of the mixed terms, valid wherever the block
                // has them.  D_x u only where D_y and D_z read it and the tile
                // points did not write it: the y halo rows of the tile's z planes
                // (NA), and the tile's y rows of the z halo planes (NB); the
                // y-halo x z-halo corners are never read.  D_y u on the z halo
                // planes.
                if (GEN) {
                    // one pass: each cell of the z halo planes (NB) gets both D_x
                    // and D_y; the y halo rows of the tile's z planes (NA) D_x
                    const int NA = TX * (2 * GW) * TZ, NB = TX * TY * (2 * GW);
                    for (int e = t; e < NA + NB; e += NT) {
                        if (e < NB) {
                            const int r = div_q(e, dTX), rz = div_q(r, dTY);
                            const int xx = e - r * TX, yy = r - rz * TY, zz = rz < GW ? rz : rz + TZ;
                            const int gi = tc.i0 + xx, gj = tc.j0 + yy, gk = tc.k0 - GW + zz;
                            const double *cu = &U[(xx + GW) + RS * (yy + GW) + PS * zz];
                            double vx = 0.0, vy = 0.0;
                            if (gi < n0 - GW && gk < n2) {
                                if (gj < n1) vx = d1(cu, 1, gi, n0, clo0, chi0, a0);
                                if (gj < n1 - GW) vy = d1(cu, RS, gj, n1, clo1, chi1, a1);
                            }
                            DX[xx + TX * (yy + GW) + DXZ * zz] = vx;
                            DY[xx + TX * yy + DYZ * zz] = vy;
                        } else {
                            const int r0 = e - NB, r = div_q(r0, dTX), zt = r / (2 * GW), ry = r - zt * (2 * GW);
                            const int xx = r0 - r * TX, yy = ry < GW ? ry : ry + TY, zz = GW + zt;
                            const int gi = tc.i0 + xx, gj = tc.j0 - GW + yy, gk = tc.k0 - GW + zz;
                            double v = 0.0;
                            if (gi < n0 - GW && gj < n1 && gk < n2)
                                v = d1(&U[(xx + GW) + RS * yy + PS * zz], 1, gi, n0, clo0, chi0, a0);
                            DX[xx + TX * yy + DXZ * zz] = v;
                        }
                    }
                } else {
                constexpr int NA = T * (2 * GW) * T, NB = T * T * (2 * GW);
                for (int e = t; e < NA + NB; e += NT) {
                    int xx, yy, zz;
                    if (e < NA) {
                        const int ry = (e / T) % (2 * GW);
                        xx = e % T;
                        yy = ry < GW ? ry : ry + T;
                        zz = GW + e / (T * 2 * GW);
                    } else {
                        const int r = e - NA, rz = r / (T * T);
                        xx = r % T;
                        yy = GW + (r / T) % T;
                        zz = rz < GW ? rz : rz + T;
                    }
                    const int gi = tc.i0 + xx, gj = tc.j0 - GW + yy, gk = tc.k0 - GW + zz;
                    double v = 0.0;
                    if (gi < n0 - GW && gj < n1 && gk < n2)
                        v = d1(&U[(xx + GW) + RS * (yy + TH * zz)], 1, gi, n0, clo0, chi0, a0);
                    DX[xx + T * (yy + TH * zz)] = v;
                }
                for (int e = t; e < T * T * (2 * GW); e += NT) {
                    const int rz = e / (T * T);
                    const int xx = e % T, yy = (e / T) % T, zz = rz < GW ? rz : rz + T;
                    const int gi = tc.i0 + xx, gj = tc.j0 + yy, gk = tc.k0 - GW + zz;
                    double v = 0.0;
                    if (gi < n0 - GW && gj < n1 - GW && gk < n2)
                        v = d1(&U[(xx + GW) + RS * ((yy + GW) + TH * zz)], RS, gj, n1, clo1, chi1, a1);
                    DY[xx + T * (yy + T * zz)] = v;
                }
                }
                __syncthreads();
                if (valid) {
                    const double *cxp = &DX[cx];
                    const double *cyp = &DY[cy];
                    const double m01 = d1(cxp, TX, pj, n1, clo1, chi1, a1);   // D_y D_x
                    const double m02 = d1(cxp, DXZ, pk, n2, clo2, chi2, a2);  // D_z D_x
                    const double m12 = d1(cyp, DYZ, pk, n2, clo2, chi2, a2);  // D_z D_y
                    if (f == 0 || f == 1) {
                        ws_put(&wsb[rs_d2(f, 1) * SS + wp], (W)m01);
                        ws_put(&wsb[rs_d2(f, 2) * SS + wp], (W)m02);
                        ws_put(&wsb[rs_d2(f, 4) * SS + wp], (W)m12);
                    } else {
                        // the shift and gt_ij enter only contracted: igt^{lm} d_l d_m f
                        // is stored, and for the shift d_j d_k beta^k is accumulated
                        SGR_EXACT
                        const double *ig = igp;
                        const double L = ((((ig[0] * h0 + 2.0 * (ig[1] * m01)) + 2.0 * (ig[2] * m02)) + ig[3] * h1) +
                                          2.0 * (ig[4] * m12)) + ig[5] * h2;
                        ws_put(&wsb[rs_lap(f) * SS + wp], (W)L);
                        if (f <= 8) {  // shift beta^k, k = f - 6: d_j d_k beta^k for j = 0, 1, 2
                            const double r0 = f == 6 ? h0 : (f == 7 ? m01 : m02);
                            const double r1 = f == 6 ? m01 : (f == 7 ? h1 : m12);
                            const double r2 = f == 6 ? m02 : (f == 7 ? m12 : h2);
                            double *gd = &GD[t];
                            if (f == 6) {
                                gd[0] = r0;
                                gd[TILE_PTS] = r1;
                                gd[2 * TILE_PTS] = r2;
                            } else if (f == 7) {
                                gd[0] = gd[0] + r0;
                                gd[TILE_PTS] = gd[TILE_PTS] + r1;
                                gd[2 * TILE_PTS] = gd[2 * TILE_PTS] + r2;
                            } else {
                                ws_put(&wsb[rs_gdb(0) * SS + wp], (W)(gd[0] + r0));
                                ws_put(&wsb[rs_gdb(1) * SS + wp], (W)(gd[TILE_PTS] + r1));
                                ws_put(&wsb[rs_gdb(2) * SS + wp], (W)(gd[2 * TILE_PTS] + r2));
                            }
                        }
                    }
                }
            }
            __syncthreads();  // every read of U (and DX, DY) for field f is done
            if (f + 1 < NFLD) {
    #pragma unroll
                for (int q = 0; q < NPF; q++)
                    if (t + NT * q < NHALO) U[lds_at(t + NT * q)] = pf[q];
            }
            __syncthreads();
        }
        if constexpr (!QUEUE) break;
        __syncthreads();  // the tile's last reads of U, DX, DY and GD are done
    }
}

// the BSSN system (gen/bssn_algebra_pw.inc, sympygr_amd/codegen/bssn_eqs.py)
#define SGR_PW_NFLD NF
#define SGR_PW_NAME k_pointwise
#define SGR_PW_ALG "gen/bssn_algebra_pw.inc"
#include "pointwise_body.h"
#undef SGR_PW_NAME
#undef SGR_PW_ALG
#undef SGR_PW_NFLD
// the CCZ4 system (gen/ccz4_algebra_pw.inc, sympygr_amd/codegen/ccz4_eqns.py)
#define SGR_PW_NFLD NF_CCZ4
#define SGR_PW_NAME k_ccz4_pointwise
#define SGR_PW_ALG "gen/ccz4_algebra_pw.inc"
#include "pointwise_body.h"
#undef SGR_PW_NAME
#undef SGR_PW_ALG
#undef SGR_PW_NFLD

// ------------------------------------------------------------ TVB dissipation
// Characteristic speeds at every node of the field arrays, ghost nodes
// included (bssn/src/rhs.cpp:984-1028 fills interior nodes only):
// lam[d span + p] = max(|-beta^d + t|, |-beta^d - t|), t = alpha sqrt(gtu_dd chi),
// gtu_dd the diagonal cofactors of gt (det gt = 1 assumed), all three 1
// where one is negative.  Nodes between blocks get values nothing reads.
__global__ __launch_bounds__(256) void k_speeds(FieldPtrs in, double *__restrict__ lam, uint64_t span)
{
    SGR_EXACT
    const uint64_t p = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (p >= span) return;
    const double g0 = in.p[12][p], g1 = in.p[13][p], g2 = in.p[14][p];
    const double g3 = in.p[15][p], g4 = in.p[16][p], g5 = in.p[17][p];
    double c[3] = {g3 * g5 - g4 * g4, g0 * g5 - g2 * g2, g0 * g3 - g1 * g1};
    if (c[0] < 0.0 || c[1] < 0.0 || c[2] < 0.0) c[0] = c[1] = c[2] = 1.0;
    const double alpha = in.p[0][p], chi = in.p[1][p];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const double t = alpha * sqrt(c[d] * chi), b = in.p[6 + d][p];
        lam[d * span + p] = larger(fabs(-b + t), fabs(-b - t));
    }
}

// Adds sigma (D_x + D_y + D_z) of TVB3 or TVB5 to the tail of every field at
// every interior point the launch's tiles own (after the stencil pass, which
// left the Kreiss-Oliger term out; the reference adds the dissipation after
// the boundary conditions, rhs.cpp:766-802, so face points get it too).  Same
// tile -> thread map as the pointwise kernel.
template <bool TVB5>
__global__ __launch_bounds__(256) void k_tvb(FieldPtrs in, const double *__restrict__ lam, uint64_t span,
                                             const BlockDesc *__restrict__ blocks,
                                             const uint32_t *__restrict__ tile_block, double *__restrict__ wsT,
                                             double sigma, uint32_t tile0)
{
    const uint32_t wg = xcd_remap(blockIdx.x, gridDim.x);
    const uint32_t tile = tile0 + (wg >> 1);
    const int l = (int)((wg & 1) * 256 + threadIdx.x);
    const BlockDesc &B = blocks[tile_block[tile]];
    const TileCoord tc = tile_coord(B, tile);
    const int TX = B.tdim[0], TY = B.tdim[1];
    if (l >= TX * TY * B.tdim[2]) return;
    const int r = div_q(l, make_div(TX)), lz = div_q(r, make_div(TY));
    const int pos[3] = {tc.i0 + l - r * TX, tc.j0 + r - lz * TY, tc.k0 + lz};
    const int nn[3] = {(int)B.n[0], (int)B.n[1], (int)B.n[2]};
    if (pos[0] >= nn[0] - GW || pos[1] >= nn[1] - GW || pos[2] >= nn[2] - GW) return;
    if (pos[0] < tc.oi || pos[1] < tc.oj || pos[2] < tc.ok) return;  // owned by the previous tile
    const int st[3] = {1, nn[0], nn[0] * nn[1]};
    const uint64_t pp = B.offset + pos[0] + (uint64_t)nn[0] * (pos[1] + (uint64_t)nn[1] * pos[2]);
    // the speeds at the taps the point's forms read (TVB3: -2..2, TVB5: -4..3)
    constexpr int OLO = TVB5 ? -4 : -2, OHI = TVB5 ? 3 : 2;
    double L[3][8], pre[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        pre[d] = TVB5 ? (d == 2 ? B.ax[2].ko_pre : 2.0 / 75.0 / B.ax[d].h) : -1.0 / 12.0 / B.ax[d].h;
#pragma unroll
        for (int o = OLO; o <= OHI; o++) {
            const int q = pos[d] + o;
            L[d][o + 4] = q >= 0 && q < nn[d] ? lam[d * span + pp + (int64_t)o * st[d]] : 0.0;
        }
    }
    const uint32_t bf = B.bflag;
    double *__restrict__ tb = wsT + ws_tile_base(wg >> 1, NF) + l;
    constexpr uint64_t SS = WS_SS;
#pragma unroll 1
    for (int f = 0; f < NF; f++) {
        const double *u = in.p[f] + pp;
        double D[3];
#pragma unroll
        for (int d = 0; d < 3; d++)
            D[d] = tvb<TVB5>(u, st[d], pos[d], nn[d], bf & (1u << (2 * d)), bf & (2u << (2 * d)), L[d], pre[d], B.ax[d]);
        {
            SGR_EXACT
            tb[rs_tail(f) * SS] = tb[rs_tail(f) * SS] + sigma * ((D[0] + D[1]) + D[2]);
        }
    }
}

hipError_t launch_speeds(const LaunchArgs &a, hipStream_t s)
{
    if (!a.lam_span) return hipSuccess;
    const uint64_t g = (a.lam_span + 255) / 256;
    if (g > 0x7FFFFFFFull) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_speeds, dim3((uint32_t)g), dim3(256), 0, s, a.in, a.lam, a.lam_span);
    return hipGetLastError();
}

// ------------------------------------------------------------------ launch
__global__ void k_queue_reset(uint32_t *queue)
{
    if (threadIdx.x < 8) queue[threadIdx.x] = 0u;
}

// Stencil grid for n tiles: one workgroup per tile, or with the tile
// queues the resident workgroups (a.grid_cap, a multiple of 8) and at least
// 8 so that every XCD queue is served.
static uint32_t stencil_grid(const LaunchArgs &a, uint32_t n)
{
    if (!a.queue) return n;
    const uint32_t g = ((n + 7) / 8) * 8;
    return g < a.grid_cap ? g : a.grid_cap;
}

// One stencil launch over n tiles from tile0 (workspace tiles from wsoff);
// with the tile queues, after a kernel that zeroes them (a kernel, not a
// memset, so that a captured graph replays it).
template <typename W, bool UPWIND, int NFLD, bool GEN>
static hipError_t launch_stencil1(const LaunchArgs &a, W *ws, uint32_t tile0, uint32_t wsoff, uint32_t n,
                                  hipEvent_t e0, hipEvent_t e1, hipStream_t s)
{
    if (a.queue) {
        hipExtLaunchKernelGGL(k_queue_reset, dim3(1), dim3(64), 0, s, e0, nullptr, 0, a.queue);
        hipExtLaunchKernelGGL((k_stencil<W, UPWIND, NFLD, GEN, true>), dim3(stencil_grid(a, n)), dim3(512), 0, s,
                              nullptr, e1, 0, a.in, a.blocks, a.tile_block, ws, a.wsT, a.params, tile0, wsoff, n,
                              a.queue);
    } else {
        hipExtLaunchKernelGGL((k_stencil<W, UPWIND, NFLD, GEN, false>), dim3(n), dim3(512), 0, s, e0, e1, 0, a.in,
                              a.blocks, a.tile_block, ws, a.wsT, a.params, tile0, wsoff, n, nullptr);
    }
    return hipGetLastError();
}

// The stencil pass of one launch: the 8^3-shaped tiles (the first a.nfast)
// with the constant-indexed kernel, the others with the general one; events
// ev[0] / ev[1] bracket the pass.
template <typename W, bool UPWIND, int NFLD>
static hipError_t launch_stencil(const LaunchArgs &a, W *ws, hipStream_t s)
{
    const uint32_t ngen = a.ntiles - a.nfast;
    if (a.nfast) {
        const hipError_t e = launch_stencil1<W, UPWIND, NFLD, false>(a, ws, a.tile0, 0u, a.nfast, a.ev[0],
                                                                     ngen ? nullptr : a.ev[1], s);
        if (e != hipSuccess || !ngen) return e;
    }
    if (ngen)
        return launch_stencil1<W, UPWIND, NFLD, true>(a, ws, a.tile0 + a.nfast, a.nfast, ngen,
                                                      a.nfast ? nullptr : a.ev[0], a.ev[1], s);
    return hipSuccess;
}

template <typename R, typename W, bool UPWIND, bool RK>
static hipError_t launch_t(const LaunchArgs &a, hipStream_t s)
{
    W *ws = (W *)a.ws;
    const dim3 g(2 * a.ntiles), b(256);
    // hipExtLaunchKernelGGL records the optional events around exactly this
    // kernel on this stream (null events: plain launch).
    if (a.ccz4) {
        hipError_t e = a.only == 2 ? hipSuccess : launch_stencil<W, UPWIND, NF_CCZ4>(a, ws, s);
        if (e != hipSuccess || a.only == 1) return e;
        hipExtLaunchKernelGGL((k_ccz4_pointwise<R, W, RK>), g, b, 0, s, a.ev[2], a.ev[3], 0, a.in, a.out, a.blocks,
                              a.tile_block, (const W *)ws, (const double *)a.wsT, a.params, a.rk, a.tile0);
        return hipGetLastError();
    }
    hipError_t e = a.only == 2 ? hipSuccess : launch_stencil<W, UPWIND, NF>(a, ws, s);
    if (e != hipSuccess || a.only == 1) return e;
    if (a.diss) {  // TVB into the tails (BSSN only)
        if (a.diss == 2)
            hipLaunchKernelGGL(k_tvb<true>, g, b, 0, s, a.in, a.lam, a.lam_span, a.blocks, a.tile_block, a.wsT,
                               a.diss_sigma, a.tile0);
        else
            hipLaunchKernelGGL(k_tvb<false>, g, b, 0, s, a.in, a.lam, a.lam_span, a.blocks, a.tile_block, a.wsT,
                               a.diss_sigma, a.tile0);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    hipExtLaunchKernelGGL((k_pointwise<R, W, RK>), g, b, 0, s, a.ev[2], a.ev[3], 0, a.in, a.out, a.blocks,
                          a.tile_block, (const W *)ws, (const double *)a.wsT, a.params, a.rk, a.tile0);
    return hipGetLastError();
}

template <typename R, typename W, bool UPWIND>
static hipError_t launch_rk(const LaunchArgs &a, hipStream_t s)
{
    return a.rk.stage >= 0 ? launch_t<R, W, UPWIND, true>(a, s) : launch_t<R, W, UPWIND, false>(a, s);
}

hipError_t launch_rhs(const LaunchArgs &a, hipStream_t s)
{
    if (a.ntiles == 0) return hipSuccess;
    if (a.mixed) return a.upwind ? launch_rk<float, float, true>(a, s) : launch_rk<float, float, false>(a, s);
    return a.upwind ? launch_rk<double, double, true>(a, s) : launch_rk<double, double, false>(a, s);
}

// Hamiltonian and momentum constraints (GR/constraints.py:161-173) at every
// interior point, in double, from the derivative workspace; with PSI4 also the
// real and imaginary parts of the Weyl scalar Psi4 (GR/constraints.py:46-155)
// at the point's coordinates pmin + i h (dendrosym/codegen.py:732).
template <typename W, bool UPWIND, bool PSI4>
__global__ __launch_bounds__(256) void k_constraints(FieldPtrs in, OutPtrs cout, const BlockDesc *__restrict__ blocks,
                                                     const uint32_t *__restrict__ tile_block, const W *__restrict__ ws,
                                                     uint32_t tile0)
{
    const uint32_t wg = xcd_remap(blockIdx.x, gridDim.x);
    const uint32_t tile = tile0 + (wg >> 1);
    const int l = (int)((wg & 1) * 256 + threadIdx.x);
    const BlockDesc &B = blocks[tile_block[tile]];
    const TileCoord tc = tile_coord8(B, tile);
    const int i = tc.i0 + (l & 7), j = tc.j0 + ((l >> 3) & 7), k = tc.k0 + (l >> 6);
    const int n0 = B.n[0], n1 = B.n[1], n2 = B.n[2];
    if (i >= n0 - GW || j >= n1 - GW || k >= n2 - GW) return;
    if (i < tc.oi || j < tc.oj || k < tc.ok) return;  // owned by the previous tile
    const uint64_t pp = B.offset + i + (uint64_t)n0 * (j + (uint64_t)n1 * k);
    const W *__restrict__ wsb = ws + ws_tile_base(wg >> 1, NSLOT) + l;
    constexpr uint64_t SS = WS_SS;
    typedef double R;
#define BSSN_U(F_) ((R)in.p[F_][pp])
#define BSSN_D1(F_, D_) ((R)ws_get(&wsb[slot_d1(F_, D_) * SS]))
#define BSSN_D2(F_, S_) ((R)ws_get(&wsb[slot_d2(F_, S_) * SS]))
#define BSSN_LAP(F_) ((R)ws_get(&wsb[slot_lap(F_) * SS]))
#define BSSN_GDB(J_) ((R)ws_get(&wsb[slot_gdb(J_) * SS]))
#define BSSN_OUT(C_, V_) cout.p[C_][pp] = (V_)
#define BSSN_BARRIER()
    if constexpr (PSI4) {
        const R px = B.pmin[0] + i * B.ax[0].h, py = B.pmin[1] + j * B.ax[1].h, pz = B.pmin[2] + k * B.ax[2].h;
#include "gen/bssn_constraints_psi4.inc"
    } else {
#include "gen/bssn_constraints.inc"
    }
#undef BSSN_U
#undef BSSN_D1
#undef BSSN_D2
#undef BSSN_OUT
#undef BSSN_LAP
#undef BSSN_BARRIER
}

template <typename W, bool UPWIND>
static hipError_t launch_constraints_t(const LaunchArgs &a, OutPtrs cout, bool psi4, hipStream_t s)
{
    W *ws = (W *)a.ws;
    if (a.derivs_threads == 512)
        hipLaunchKernelGGL((k_derivs<W, UPWIND, 512>), dim3(a.ntiles), dim3(512), 0, s, a.in, a.blocks, a.tile_block, ws,
                           a.wsK, true, a.tile0);
    else
        hipLaunchKernelGGL((k_derivs<W, UPWIND, 256>), dim3(a.ntiles), dim3(256), 0, s, a.in, a.blocks, a.tile_block, ws,
                           a.wsK, true, a.tile0);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (psi4)
        hipLaunchKernelGGL((k_constraints<W, UPWIND, true>), dim3(2 * a.ntiles), dim3(256), 0, s, a.in, cout, a.blocks,
                           a.tile_block, (const W *)ws, a.tile0);
    else
        hipLaunchKernelGGL((k_constraints<W, UPWIND, false>), dim3(2 * a.ntiles), dim3(256), 0, s, a.in, cout,
                           a.blocks, a.tile_block, (const W *)ws, a.tile0);
    return hipGetLastError();
}

hipError_t launch_constraints(const LaunchArgs &a, OutPtrs cout, bool psi4, hipStream_t s)
{
    if (a.ntiles == 0) return hipSuccess;
    if (a.mixed)
        return a.upwind ? launch_constraints_t<float, true>(a, cout, psi4, s)
                        : launch_constraints_t<float, false>(a, cout, psi4, s);
    return a.upwind ? launch_constraints_t<double, true>(a, cout, psi4, s)
                    : launch_constraints_t<double, false>(a, cout, psi4, s);
}

// Taps of one axis for target node `idx`: returns the first source node and
// fills NT weights.  NT == 1: the node coincides with source node q.  NT == 6:
// 6-point Lagrange over p0..p0+5, p0 = clamp(q - 2, 3, sm - 9), at
// t = (q - p0) + rem / hs; a node that happens to coincide gets the unit
// weight vector.  The position lies in the source's owned box, so
// q is in [2, sm - 4] and every tap is an interior node.
template <int NT>
__device__ inline int ghost_axis(int idx, int rel, int ht, int hs, double inv_hs, int sm, double *w)
{
    SGR_EXACT
    const int lat = rel + idx * ht;
    const int q = lat / hs, rem = lat - q * hs;
    if (NT == 1) {
        w[0] = 1.0;
        return q;
    }
    const int p0 = min(max(q - 2, 3), sm - 9);
    if (rem == 0) {
#pragma unroll
        for (int j = 0; j < NT; j++) w[j] = (j == q - p0) ? 1.0 : 0.0;
        return p0;
    }
    constexpr double C[6] = {1.0 / -120.0, 1.0 / 24.0, 1.0 / -12.0, 1.0 / 12.0, 1.0 / -24.0, 1.0 / 120.0};
    const double t = (double)(q - p0) + (double)rem * inv_hs;
#pragma unroll
    for (int j = 0; j < NT; j++) {
        double N = 1.0;
#pragma unroll
        for (int k = 0; k < NT; k++)
            if (k != j) N = N * (t - (double)k);
        w[j] = N * C[j];
    }
    return p0;
}

// Mixed-radix counter over (i0, i1, i2), i0 fastest, advanced by a fixed
// stride: replaces two integer divisions per item in the box loops.
struct Idx3 {
    int i0, i1, i2, s0, s1, s2, n0, n1;
    __device__ Idx3(int lin, int step, int n0_, int n1_) : n0(n0_), n1(n1_)
    {
        i0 = lin % n0;
        i1 = (lin / n0) % n1;
        i2 = lin / (n0 * n1);
        s0 = step % n0;
        s1 = (step / n0) % n1;
        s2 = step / (n0 * n1);
    }
    __device__ void next()
    {
        i0 += s0;
        int c = i0 >= n0;
        i0 -= c ? n0 : 0;
        i1 += s1 + c;
        c = i1 >= n1;
        i1 -= c ? n1 : 0;
        i2 += s2 + c;
    }
};

// Separable refresh of one interpolating box, field by field through LDS:
//   T1[k][j][ix] = sum_a wx_a(ix) u(px(ix)+a, row_y j, row_z k)
//   T2[k][iy][ix] = sum_b wy_b(iy) T1[k][ry0(iy)+b][ix]
//   u(target)     = sum_c wz_c(iz) T2[rz0(iz)+c][iy][ix]
// Each value is formed by the same operations in the same order as the
// per-node sum_c wz_c (sum_b wy_b (sum_a wx_a u)) with contraction off, so
// results are bit-identical to it, while the loads per node fall from
// NX*NY*NZ to NX*Ry*Rz/(ey*ez).  Sources are interior nodes and targets
// ghost nodes, so boxes never read what another box writes.
template <int NX, int NY, int NZ>
__global__ __launch_bounds__(256) void k_ghost_sep(OutPtrs u, int nf, const GhostPatch *__restrict__ patches,
                                                   const GhostBox *__restrict__ boxes)
{
    SGR_EXACT
    __shared__ double T[GHOST_LDS];
    __shared__ double wxs[GHOST_BOX_MAX * NX], wys[GHOST_BOX_MAX * NY], wzs[GHOST_BOX_MAX * NZ];
    __shared__ int px[GHOST_BOX_MAX], ry0[GHOST_BOX_MAX], rz0[GHOST_BOX_MAX];
    __shared__ int rowy[GHOST_ROWS_MAX], rowz[GHOST_ROWS_MAX];  // source y (z) node of each row
    const GhostBox bx = boxes[blockIdx.x];
    const GhostPatch g = patches[bx.patch];
    const int ex = bx.e[0], ey = bx.e[1], ez = bx.e[2];
    const int Ry = ghost_rows(bx.lo[1], ey, g.rel[1], g.ht[1], g.hs[1], g.sm[1], NY);
    const int Rz = ghost_rows(bx.lo[2], ez, g.rel[2], g.ht[2], g.hs[2], g.sm[2], NZ);
    const int n1 = ex * Ry * Rz, n2 = ex * ey * Rz, n3 = ex * ey * ez;
    // the host splits patches into boxes within these limits
    if (ex > GHOST_BOX_MAX || ey > GHOST_BOX_MAX || ez > GHOST_BOX_MAX || Ry > GHOST_ROWS_MAX ||
        Rz > GHOST_ROWS_MAX || n1 + n2 > GHOST_LDS)
        return;
    const int t = threadIdx.x;
    const int by0 = ghost_p0(bx.lo[1], g.rel[1], g.ht[1], g.hs[1], g.sm[1], NY);
    const int bz0 = ghost_p0(bx.lo[2], g.rel[2], g.ht[2], g.hs[2], g.sm[2], NZ);
    if (t < ex) {
        double w[NX];
        px[t] = ghost_axis<NX>(bx.lo[0] + t, g.rel[0], g.ht[0], g.hs[0], g.inv_hs[0], g.sm[0], w);
#pragma unroll
        for (int a = 0; a < NX; a++) wxs[t * NX + a] = w[a];
    } else if (t >= 64 && t - 64 < ey) {
        const int iy = t - 64;
        double w[NY];
        const int p = ghost_axis<NY>(bx.lo[1] + iy, g.rel[1], g.ht[1], g.hs[1], g.inv_hs[1], g.sm[1], w);
        ry0[iy] = NY == 1 ? iy : p - by0;
#pragma unroll
        for (int b = 0; b < NY; b++) wys[iy * NY + b] = w[b];
    } else if (t >= 128 && t - 128 < ez) {
        const int iz = t - 128;
        double w[NZ];
        const int p = ghost_axis<NZ>(bx.lo[2] + iz, g.rel[2], g.ht[2], g.hs[2], g.inv_hs[2], g.sm[2], w);
        rz0[iz] = NZ == 1 ? iz : p - bz0;
#pragma unroll
        for (int c = 0; c < NZ; c++) wzs[iz * NZ + c] = w[c];
    }
    for (int j = t; j < Ry; j += 256)
        rowy[j] = NY == 1 ? ghost_p0(bx.lo[1] + j, g.rel[1], g.ht[1], g.hs[1], g.sm[1], 1) : by0 + j;
    for (int k = t; k < Rz; k += 256)
        rowz[k] = NZ == 1 ? ghost_p0(bx.lo[2] + k, g.rel[2], g.ht[2], g.hs[2], g.sm[2], 1) : bz0 + k;
    __syncthreads();
    const uint64_t dy = g.dn0, dz = (uint64_t)g.dn0 * g.dn1;
    const uint64_t dst0 = g.dst_off + bx.lo[0] + dy * bx.lo[1] + dz * bx.lo[2];
    double *T1 = T, *T2 = T + n1;
    const uint64_t sy = g.sn0, sz = (uint64_t)g.sn0 * g.sn1;
#pragma unroll 1
    for (int f = 0; f < nf; f++) {
        const double *__restrict__ v = u.p[f] + g.src_off;
        {
            Idx3 it(t, 256, ex, Ry);
            for (int i = t; i < n1; i += 256, it.next()) {
                const double *r = v + px[it.i0] + sy * rowy[it.i1] + sz * rowz[it.i2];
                double row = 0.0;
#pragma unroll
                for (int a = 0; a < NX; a++) row = row + wxs[it.i0 * NX + a] * r[a];
                T1[i] = row;
            }
        }
        __syncthreads();
        {
            Idx3 it(t, 256, ex, ey);
            for (int i = t; i < n2; i += 256, it.next()) {
                const double *r = T1 + (it.i2 * Ry + ry0[it.i1]) * ex + it.i0;
                double plane = 0.0;
#pragma unroll
                for (int b = 0; b < NY; b++) plane = plane + wys[it.i1 * NY + b] * r[b * ex];
                T2[i] = plane;
            }
        }
        __syncthreads();
        {
            double *__restrict__ o = u.p[f] + dst0;
            Idx3 it(t, 256, ex, ey);
            for (int i = t; i < n3; i += 256, it.next()) {
                const double *r = T2 + (rz0[it.i2] * ey + it.i1) * ex + it.i0;
                double out = 0.0;
#pragma unroll
                for (int c = 0; c < NZ; c++) out = out + wzs[it.i2 * NZ + c] * r[c * ey * ex];
                o[it.i0 + dy * it.i1 + dz * it.i2] = out;
            }
        }
        // the next field's first pass writes only T1, whose readers finished
        // before the barrier above; T2 is rewritten after its next barrier
    }
}

// Boxes whose nodes coincide with source nodes on every axis: copies, written
// as the per-node form evaluates them (0.0 + 1.0*u, which maps -0.0 to +0.0).
__global__ __launch_bounds__(256) void k_ghost_copy(OutPtrs u, int nf, const GhostPatch *__restrict__ patches,
                                                    const GhostBox *__restrict__ boxes)
{
    SGR_EXACT
    const GhostBox bx = boxes[blockIdx.x];
    const GhostPatch g = patches[bx.patch];
    const int ex = bx.e[0], ey = bx.e[1], n = ex * ey * bx.e[2];
    const uint64_t sy = g.sn0, sz = (uint64_t)g.sn0 * g.sn1;
    const uint64_t dy = g.dn0, dz = (uint64_t)g.dn0 * g.dn1;
    Idx3 it(threadIdx.x, 256, ex, ey);
    for (int i = threadIdx.x; i < n; i += 256, it.next()) {
        const int x = bx.lo[0] + it.i0, y = bx.lo[1] + it.i1, z = bx.lo[2] + it.i2;
        const int ax = ghost_p0(x, g.rel[0], g.ht[0], g.hs[0], g.sm[0], 1);
        const int ay = ghost_p0(y, g.rel[1], g.ht[1], g.hs[1], g.sm[1], 1);
        const int az = ghost_p0(z, g.rel[2], g.ht[2], g.hs[2], g.sm[2], 1);
        const uint64_t src = g.src_off + ax + sy * ay + sz * az;
        const uint64_t dst = g.dst_off + x + dy * y + dz * z;
#pragma unroll 4
        for (int f = 0; f < nf; f++) u.p[f][dst] = 0.0 + 1.0 * u.p[f][src];
    }
}

template <int CLS>
static void launch_ghost_class(OutPtrs u, int nf, const GhostPatch *patches, const GhostBox *boxes,
                               const uint32_t *bstart, hipStream_t s)
{
    const uint32_t n = bstart[CLS + 1] - bstart[CLS];
    if (n == 0) return;
    constexpr int NX = (CLS & 1) ? 1 : 6, NY = (CLS & 2) ? 1 : 6, NZ = (CLS & 4) ? 1 : 6;
    if constexpr (CLS == GHOST_CLASSES - 1)
        hipLaunchKernelGGL(k_ghost_copy, dim3(n), dim3(256), 0, s, u, nf, patches, boxes + bstart[CLS]);
    else
        hipLaunchKernelGGL((k_ghost_sep<NX, NY, NZ>), dim3(n), dim3(256), 0, s, u, nf, patches, boxes + bstart[CLS]);
}

hipError_t launch_ghost_fill(OutPtrs u, int nf, const GhostPatch *patches, const GhostBox *boxes,
                             const uint32_t bstart[GHOST_CLASSES + 1], hipStream_t s)
{
    launch_ghost_class<0>(u, nf, patches, boxes, bstart, s);
    launch_ghost_class<1>(u, nf, patches, boxes, bstart, s);
    launch_ghost_class<2>(u, nf, patches, boxes, bstart, s);
    launch_ghost_class<3>(u, nf, patches, boxes, bstart, s);
    launch_ghost_class<4>(u, nf, patches, boxes, bstart, s);
    launch_ghost_class<5>(u, nf, patches, boxes, bstart, s);
    launch_ghost_class<6>(u, nf, patches, boxes, bstart, s);
    launch_ghost_class<7>(u, nf, patches, boxes, bstart, s);
    return hipGetLastError();
}

}  // namespace sgr
