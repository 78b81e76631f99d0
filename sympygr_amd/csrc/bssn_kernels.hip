// BSSN RHS kernels for gfx950 (CDNA4).
//
// Two passes over 8x8x8 interior tiles of every block (tile-major workspace):
//
//   k_derivs     one workgroup per tile.  Field by field, the 14^3 halo tile
//                is staged in LDS; every thread evaluates the first, second,
//                advective and Kreiss-Oliger stencils of its two points from
//                LDS.  The inner first derivatives of the mixed terms
//                (D_x u on the tile's y/z halo, D_y u on its z halo) are
//                staged in LDS as well, so D_y D_x, D_z D_x and D_z D_y are
//                the same compositions as the reference's sweeps
//                (bssn/scripts/bssnrhs_derivs.h:7-9).  Results go to the
//                derivative workspace (double, or float in mixed mode) and
//                the KO sums (always double).
//
//   k_pointwise  one thread per interior point: reads its fields and
//                derivatives, runs the generated algebra in R (double or
//                float), applies the Sommerfeld condition on flagged faces,
//                adds sigma * KO in double and writes the RHS.
//
// Both kernels read the block table and a tile->block map, so one launch
// covers any mix of block sizes.
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
};

__device__ __forceinline__ TileCoord tile_coord(const BlockDesc &B, uint32_t tile)
{
    TileCoord c;
    c.lt = (int)(tile - B.tile_base);
    const int tx = c.lt % B.ntile[0];
    const int ty = (c.lt / B.ntile[0]) % B.ntile[1];
    const int tz = c.lt / (B.ntile[0] * B.ntile[1]);
    c.i0 = GW + T * tx;
    c.j0 = GW + T * ty;
    c.k0 = GW + T * tz;
    return c;
}

template <typename W, bool UPWIND>
__global__ __launch_bounds__(256) void k_derivs(FieldPtrs in, const BlockDesc *__restrict__ blocks,
                                                const uint32_t *__restrict__ tile_block, W *__restrict__ ws,
                                                double *__restrict__ wsK, uint64_t NP)
{
    __shared__ double U[TH * TH * TH];   // x + 14 (y + 14 z), halo coordinates
    __shared__ double DX[T * TH * TH];   // D_x u: x in tile, y and z with halo
    __shared__ double DY[T * T * TH];    // D_y u: x, y in tile, z with halo

    const uint32_t tile = blockIdx.x;
    const BlockDesc &B = blocks[tile_block[tile]];
    const TileCoord tc = tile_coord(B, tile);
    const int n0 = B.n[0], n1 = B.n[1], n2 = B.n[2];
    const bool lo0 = B.bflag & 1u, hi0 = B.bflag & 2u, lo1 = B.bflag & 4u, hi1 = B.bflag & 8u;
    const bool lo2 = B.bflag & 16u, hi2 = B.bflag & 32u;
    const AxisConst a0 = B.ax[0], a1 = B.ax[1], a2 = B.ax[2];

    const int t = threadIdx.x;
    const int lx = t & 7, ly = (t >> 3) & 7, lz = t >> 6;  // lz in [0,4); points lz and lz+4
    const int pi = tc.i0 + lx, pj = tc.j0 + ly;
    bool valid[2];
    int pk[2];
    uint64_t wp[2];
    double beta[2][3];
#pragma unroll
    for (int p = 0; p < 2; p++) {
        pk[p] = tc.k0 + lz + 4 * p;
        valid[p] = pi < n0 - GW && pj < n1 - GW && pk[p] < n2 - GW;
        wp[p] = B.ws_base + (uint64_t)tc.lt * TILE_PTS + (lx + T * (ly + T * (lz + 4 * p)));
        const uint64_t gp = B.offset + pi + (uint64_t)n0 * (pj + (uint64_t)n1 * pk[p]);
#pragma unroll
        for (int d = 0; d < 3; d++) beta[p][d] = (UPWIND && valid[p]) ? in.p[6 + d][gp] : 0.0;
    }

    for (int f = 0; f < NF; f++) {
        __syncthreads();
        const double *__restrict__ src = in.p[f] + B.offset;
        for (int e = t; e < TH * TH * TH; e += 256) {
            const int ex = e % TH, ey = (e / TH) % TH, ez = e / (TH * TH);
            const int gi = tc.i0 - GW + ex, gj = tc.j0 - GW + ey, gk = tc.k0 - GW + ez;
            U[e] = (gi < n0 && gj < n1 && gk < n2) ? src[gi + (uint64_t)n0 * (gj + (uint64_t)n1 * gk)] : 0.0;
        }
        __syncthreads();

#pragma unroll
        for (int p = 0; p < 2; p++) {
            if (!valid[p]) continue;
            const double *c = &U[(lx + GW) + TH * ((ly + GW) + TH * (lz + 4 * p + GW))];
            ws[slot_d1(f, 0) * NP + wp[p]] = (W)d1(c, 1, pi, n0, lo0, hi0, a0);
            ws[slot_d1(f, 1) * NP + wp[p]] = (W)d1(c, TH, pj, n1, lo1, hi1, a1);
            ws[slot_d1(f, 2) * NP + wp[p]] = (W)d1(c, TH * TH, pk[p], n2, lo2, hi2, a2);
            if (UPWIND) {
                ws[slot_ad(f, 0) * NP + wp[p]] = (W)adv(c, 1, pi, n0, lo0, hi0, beta[p][0], a0);
                ws[slot_ad(f, 1) * NP + wp[p]] = (W)adv(c, TH, pj, n1, lo1, hi1, beta[p][1], a1);
                ws[slot_ad(f, 2) * NP + wp[p]] = (W)adv(c, TH * TH, pk[p], n2, lo2, hi2, beta[p][2], a2);
            }
            {
                SGR_EXACT
                const double kx = ko(c, 1, pi, n0, lo0, hi0, a0);
                const double ky = ko(c, TH, pj, n1, lo1, hi1, a1);
                const double kz = ko(c, TH * TH, pk[p], n2, lo2, hi2, a2);
                wsK[(uint64_t)f * NP + wp[p]] = kx + ky + kz;
            }
        }

        if (d2q(f) < 0) continue;  // uniform over the workgroup

#pragma unroll
        for (int p = 0; p < 2; p++) {
            if (!valid[p]) continue;
            const double *c = &U[(lx + GW) + TH * ((ly + GW) + TH * (lz + 4 * p + GW))];
            ws[slot_d2(f, 0) * NP + wp[p]] = (W)d2(c, 1, pi, n0, lo0, hi0, a0);
            ws[slot_d2(f, 3) * NP + wp[p]] = (W)d2(c, TH, pj, n1, lo1, hi1, a1);
            ws[slot_d2(f, 5) * NP + wp[p]] = (W)d2(c, TH * TH, pk[p], n2, lo2, hi2, a2);
        }
        // inner derivatives of the mixed terms, valid wherever the block has them
        for (int e = t; e < T * TH * TH; e += 256) {
            const int x = e % T, y = (e / T) % TH, z = e / (T * TH);
            const int gi = tc.i0 + x, gj = tc.j0 - GW + y, gk = tc.k0 - GW + z;
            double v = 0.0;
            if (gi < n0 - GW && gj < n1 && gk < n2) v = d1(&U[(x + GW) + TH * (y + TH * z)], 1, gi, n0, lo0, hi0, a0);
            DX[e] = v;
        }
        for (int e = t; e < T * T * TH; e += 256) {
            const int x = e % T, y = (e / T) % T, z = e / (T * T);
            const int gi = tc.i0 + x, gj = tc.j0 + y, gk = tc.k0 - GW + z;
            double v = 0.0;
            if (gi < n0 - GW && gj < n1 - GW && gk < n2)
                v = d1(&U[(x + GW) + TH * ((y + GW) + TH * z)], TH, gj, n1, lo1, hi1, a1);
            DY[e] = v;
        }
        __syncthreads();
#pragma unroll
        for (int p = 0; p < 2; p++) {
            if (!valid[p]) continue;
            const double *cx = &DX[lx + T * ((ly + GW) + TH * (lz + 4 * p + GW))];
            const double *cy = &DY[lx + T * (ly + T * (lz + 4 * p + GW))];
            ws[slot_d2(f, 1) * NP + wp[p]] = (W)d1(cx, T, pj, n1, lo1, hi1, a1);
            ws[slot_d2(f, 2) * NP + wp[p]] = (W)d1(cx, T * TH, pk[p], n2, lo2, hi2, a2);
            ws[slot_d2(f, 4) * NP + wp[p]] = (W)d1(cy, T * T, pk[p], n2, lo2, hi2, a2);
        }
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

template <typename R, typename W, bool UPWIND>
__global__ __launch_bounds__(256) void k_pointwise(FieldPtrs in, OutPtrs out, const BlockDesc *__restrict__ blocks,
                                                   const uint32_t *__restrict__ tile_block,
                                                   const W *__restrict__ ws, const double *__restrict__ wsK,
                                                   uint64_t NP, Params P)
{
    const uint32_t tile = blockIdx.x >> 1;
    const int l = (int)((blockIdx.x & 1) * 256 + threadIdx.x);
    const BlockDesc &B = blocks[tile_block[tile]];
    const TileCoord tc = tile_coord(B, tile);
    const int lx = l & 7, ly = (l >> 3) & 7, lz = l >> 6;
    const int i = tc.i0 + lx, j = tc.j0 + ly, k = tc.k0 + lz;
    const int n0 = B.n[0], n1 = B.n[1], n2 = B.n[2];
    if (i >= n0 - GW || j >= n1 - GW || k >= n2 - GW) return;
    const uint64_t wp = B.ws_base + (uint64_t)tc.lt * TILE_PTS + l;
    const uint64_t pp = B.offset + i + (uint64_t)n0 * (j + (uint64_t)n1 * k);

    // coordinates and eta exactly as bssn/src/rhs.cpp:206-219
    const double x = coord(B.pmin[0], i, B.ax[0].h);
    const double y = coord(B.pmin[1], j, B.ax[1].h);
    const double z = coord(B.pmin[2], k, B.ax[2].h);
    double eta_d = P.eta_const;
    {
        const double r_coord = radius(x, y, z);
        if (r_coord >= P.eta_r0) eta_d *= pow(P.eta_r0 / r_coord, P.eta_damping_exp);
    }
    const uint32_t bf = B.bflag;
    const bool face = bf && (((bf & 1u) && i == GW) || ((bf & 2u) && i == n0 - GW - 1) ||
                             ((bf & 4u) && j == GW) || ((bf & 8u) && j == n1 - GW - 1) ||
                             ((bf & 16u) && k == GW) || ((bf & 32u) && k == n2 - GW - 1));
    const double inv_r = face ? 1.0 / radius(x, y, z) : 0.0;

    const R lambda0 = (R)P.lambda[0], lambda1 = (R)P.lambda[1], lambda2 = (R)P.lambda[2], lambda3 = (R)P.lambda[3];
    const R lambda_f0 = (R)P.lambda_f[0], lambda_f1 = (R)P.lambda_f[1];
    const R eta = (R)eta_d;
    const double sigma = P.ko_sigma;

#define BSSN_U(F_) ((R)in.p[F_][pp])
#define BSSN_D1(F_, D_) ((R)ws[slot_d1(F_, D_) * NP + wp])
#define BSSN_D2(F_, S_) ((R)ws[slot_d2(F_, S_) * NP + wp])
#define BSSN_AD(F_, D_) ((R)ws[(UPWIND ? slot_ad(F_, D_) : slot_d1(F_, D_)) * NP + wp])
#define BSSN_OUT(F_, V_)                                                                                      \
    do {                                                                                                    \
        double val_ = (double)(V_);                                                                          \
        if (face)                                                                                           \
            val_ = sommerfeld(x, y, z, inv_r, (double)ws[slot_d1(F_, 0) * NP + wp],                          \
                              (double)ws[slot_d1(F_, 1) * NP + wp], (double)ws[slot_d1(F_, 2) * NP + wp],     \
                              in.p[F_][pp], bc_falloff(F_), bc_asym(F_));                                      \
        out.p[F_][pp] = ko_add(val_, sigma, wsK[(uint64_t)(F_) * NP + wp]);                                   \
    } while (0)
#define BSSN_BARRIER() SGR_ALGEBRA_BARRIER
#include "gen/bssn_algebra.inc"
#undef BSSN_BARRIER
#undef BSSN_U
#undef BSSN_D1
#undef BSSN_D2
#undef BSSN_AD
#undef BSSN_OUT
}

// ------------------------------------------------------------------ launch
template <typename R, typename W, bool UPWIND>
static hipError_t launch_t(const LaunchArgs &a, hipStream_t s)
{
    W *ws = (W *)a.ws;
    // hipExtLaunchKernelGGL records the optional events around exactly this
    // kernel on this stream (null events: plain launch).
    hipExtLaunchKernelGGL((k_derivs<W, UPWIND>), dim3(a.ntiles), dim3(256), 0, s, a.ev[0], a.ev[1], 0, a.in,
                          a.blocks, a.tile_block, ws, a.wsK, a.np);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipExtLaunchKernelGGL((k_pointwise<R, W, UPWIND>), dim3(2 * a.ntiles), dim3(256), 0, s, a.ev[2], a.ev[3], 0,
                          a.in, a.out, a.blocks, a.tile_block, (const W *)ws, (const double *)a.wsK, a.np, a.params);
    return hipGetLastError();
}

hipError_t launch_rhs(const LaunchArgs &a, hipStream_t s)
{
    if (a.ntiles == 0) return hipSuccess;
    if (a.mixed) return a.upwind ? launch_t<float, float, true>(a, s) : launch_t<float, float, false>(a, s);
    return a.upwind ? launch_t<double, double, true>(a, s) : launch_t<double, double, false>(a, s);
}

}  // namespace sgr
