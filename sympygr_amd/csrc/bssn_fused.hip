// Fused single-pass BSSN RHS kernel for gfx950 (CDNA4).
//
// One workgroup (256 threads, one interior point each) owns an 8x8x4 tile of
// a block.  The 24 fields are visited in a fixed order (the field-staged
// algebra in gen/bssn_staged.inc): while field f's 14x14x10 halo tile sits in
// LDS, every thread evaluates the derivatives of f at its point -- first,
// second, mixed (through LDS-staged inner derivatives D_x f and D_y f, the
// same compositions as bssn/scripts/bssnrhs_derivs.h:7-9), advective and
// Kreiss-Oliger -- and immediately runs the part of the algebra those values
// complete.  The next field's tile is fetched from HBM into registers at the
// start of each stage and written to the other LDS buffer at its end, so the
// global-memory latency overlaps the stencil and algebra work.  Nothing but
// the 24 input fields is read and nothing but the 24 RHS values is written.
#include <hip/hip_ext.h>

#include "bssn_kernels.h"
#include "stencils.h"

namespace sgr {

namespace fused {
constexpr int TX = 8, TY = 8, TZ = 4;
constexpr int HX = TX + 2 * GW, HY = TY + 2 * GW, HZ = TZ + 2 * GW;
constexpr int NH = HX * HY * HZ;           // 1960 halo points
constexpr int NT = 256;                    // threads = tile points
constexpr int NPF = (NH + NT - 1) / NT;    // prefetch registers per thread
}  // namespace fused

// First padded coordinate of tile t along an axis of padded size n: tiles of
// width w from 3, the last one shifted back so it ends at the interior's end
// (it then overlaps its neighbour; shared points are computed twice with the
// same result).  Blocks narrower than w keep the origin at 3.
__device__ __forceinline__ int tile_origin(int t, int w, int n)
{
    const int o = GW + w * t;
    const int last = n - GW - w;
    return o < last ? o : (last > GW ? last : GW);
}

__device__ __forceinline__ double sommerfeld_f(double x, double y, double z, double inv_r, double gx, double gy,
                                               double gz, double fval, double falloff, double asym)
{
    SGR_EXACT
    return -inv_r * (x * gx + y * gy + z * gz + falloff * (fval - asym));
}

__device__ __forceinline__ double ko_add_f(double v, double sigma, double k)
{
    SGR_EXACT
    return v + sigma * k;
}

__device__ __forceinline__ double coord_f(double pmin, int i, double h)
{
    SGR_EXACT
    return pmin + i * h;
}

__device__ __forceinline__ double radius_f(double x, double y, double z)
{
    SGR_EXACT
    return sqrt(x * x + y * y + z * z);
}

// advection of the current field: (beta^0 ad_0 + beta^1 ad_1) + beta^2 ad_2
__device__ inline double fused_adv(const double *cu, int i, int j, int k, int n0, int n1, int n2, bool lo0, bool hi0,
                                   bool lo1, bool hi1, bool lo2, bool hi2, const AxisConst &a0, const AxisConst &a1,
                                   const AxisConst &a2, double bv0, double bv1, double bv2, bool upwind)
{
    SGR_EXACT
    double q0, q1, q2;
    if (upwind) {
        q0 = adv(cu, 1, i, n0, lo0, hi0, bv0 > 0.0 ? 1.0 : 0.0, a0);
        q1 = adv(cu, fused::HX, j, n1, lo1, hi1, bv1 > 0.0 ? 1.0 : 0.0, a1);
        q2 = adv(cu, fused::HX * fused::HY, k, n2, lo2, hi2, bv2 > 0.0 ? 1.0 : 0.0, a2);
    } else {
        q0 = d1(cu, 1, i, n0, lo0, hi0, a0);
        q1 = d1(cu, fused::HX, j, n1, lo1, hi1, a1);
        q2 = d1(cu, fused::HX * fused::HY, k, n2, lo2, hi2, a2);
    }
    return (bv0 * q0 + bv1 * q1) + bv2 * q2;
}

template <typename R, bool UPWIND>
__global__ __launch_bounds__(256) void k_fused(FieldPtrs in, OutPtrs out, const BlockDesc *__restrict__ blocks,
                                               const uint32_t *__restrict__ tile_block, Params P)
{
    using namespace fused;
    __shared__ double U[2][NH];
    __shared__ double DX[TX * HY * HZ];  // D_x f: x in tile, y and z with halo
    __shared__ double DY[TX * TY * HZ];  // D_y f: x, y in tile, z with halo

    const uint32_t tile = xcd_remap(blockIdx.x, gridDim.x);
    const BlockDesc &B = blocks[tile_block[tile]];
    const int n0 = B.n[0], n1 = B.n[1], n2 = B.n[2];
    const int lt = (int)(tile - B.tile_base);
    const int i0 = tile_origin(lt % B.ntile[0], TX, n0);
    const int j0 = tile_origin((lt / B.ntile[0]) % B.ntile[1], TY, n1);
    const int k0 = tile_origin(lt / (B.ntile[0] * B.ntile[1]), TZ, n2);
    const bool lo0 = B.bflag & 1u, hi0 = B.bflag & 2u, lo1 = B.bflag & 4u, hi1 = B.bflag & 8u;
    const bool lo2 = B.bflag & 16u, hi2 = B.bflag & 32u;
    const AxisConst a0 = B.ax[0], a1 = B.ax[1], a2 = B.ax[2];

    const int t = threadIdx.x;
    const int lx = t & 7, ly = (t >> 3) & 7, lz = t >> 6;
    const int i = i0 + lx, j = j0 + ly, k = k0 + lz;
    const bool valid = i < n0 - GW && j < n1 - GW && k < n2 - GW;
    const uint64_t pp = B.offset + i + (uint64_t)n0 * (j + (uint64_t)n1 * k);
    const int c = (lx + GW) + HX * ((ly + GW) + HY * (lz + GW));  // my point in the halo tile

    // coordinates, eta(r) and the face flag exactly as bssn/src/rhs.cpp:206-219, 833-977
    const double x = coord_f(B.pmin[0], i, a0.h);
    const double y = coord_f(B.pmin[1], j, a1.h);
    const double z = coord_f(B.pmin[2], k, a2.h);
    double eta_d = P.eta_const;
    {
        const double r_coord = radius_f(x, y, z);
        if (r_coord >= P.eta_r0) eta_d *= pow(P.eta_r0 / r_coord, P.eta_damping_exp);
    }
    const uint32_t bf = B.bflag;
    const bool face = valid && bf &&
                      (((bf & 1u) && i == GW) || ((bf & 2u) && i == n0 - GW - 1) || ((bf & 4u) && j == GW) ||
                       ((bf & 8u) && j == n1 - GW - 1) || ((bf & 16u) && k == GW) || ((bf & 32u) && k == n2 - GW - 1));
    const double inv_r = face ? 1.0 / radius_f(x, y, z) : 0.0;
    const R lambda0 = (R)P.lambda[0], lambda1 = (R)P.lambda[1], lambda2 = (R)P.lambda[2], lambda3 = (R)P.lambda[3];
    const R lambda_f0 = (R)P.lambda_f[0], lambda_f1 = (R)P.lambda_f[1];
    const R eta = (R)eta_d;
    const double sigma = P.ko_sigma;
    // shift at the point (advection; upwind direction beta^d > 0, bssn/src/derivs.cpp:518)
    const double bv0 = valid ? in.p[6][pp] : 0.0, bv1 = valid ? in.p[7][pp] : 0.0, bv2 = valid ? in.p[8][pp] : 0.0;

    // halo element e of the tile -> global offset in the block (or -1)
    auto halo_src = [&](int e) -> int32_t {
        const int ex = e % HX, ey = (e / HX) % HY, ez = e / (HX * HY);
        const int gi = i0 - GW + ex, gj = j0 - GW + ey, gk = k0 - GW + ez;
        if (e >= NH || gi >= n0 || gj >= n1 || gk >= n2) return -1;
        return gi + n0 * (gj + n1 * gk);
    };
    int32_t src[NPF];
#pragma unroll
    for (int q = 0; q < NPF; q++) src[q] = halo_src(t + NT * q);

    const double *cur = U[0];

#define BSSN_U(F_) ((R)in.p[F_][pp])

#define BSSN_FIRST_FIELD(F_)                                                                                   \
    {                                                                                                          \
        const double *s_ = in.p[F_] + B.offset;                                                               \
        _Pragma("unroll") for (int q = 0; q < NPF; q++) if (t + NT * q < NH) U[0][t + NT * q] =               \
            src[q] >= 0 ? s_[src[q]] : 0.0;                                                                    \
        __syncthreads();                                                                                       \
    }

#define BSSN_STAGE_BEGIN(F_, S_, D2_, FN_)                                                                    \
    cur = U[(S_) & 1];                                                                                         \
    double pf_##S_[NPF];                                                                                       \
    if ((FN_) >= 0) {                                                                                          \
        const double *s_ = in.p[(FN_) >= 0 ? (FN_) : 0] + B.offset;                                           \
        _Pragma("unroll") for (int q = 0; q < NPF; q++) pf_##S_[q] = src[q] >= 0 ? s_[src[q]] : 0.0;          \
    }                                                                                                          \
    if (D2_) {                                                                                                 \
        for (int e = t; e < TX * HY * HZ; e += NT) {                                                          \
            const int xx = e % TX, yy = (e / TX) % HY, zz = e / (TX * HY);                                    \
            const int gi = i0 + xx, gj = j0 - GW + yy, gk = k0 - GW + zz;                                     \
            DX[e] = (gi < n0 - GW && gj < n1 && gk < n2)                                                      \
                        ? d1(&cur[(xx + GW) + HX * (yy + HY * zz)], 1, gi, n0, lo0, hi0, a0)                  \
                        : 0.0;                                                                                 \
        }                                                                                                      \
        for (int e = t; e < TX * TY * HZ; e += NT) {                                                          \
            const int xx = e % TX, yy = (e / TX) % TY, zz = e / (TX * TY);                                    \
            const int gi = i0 + xx, gj = j0 + yy, gk = k0 - GW + zz;                                          \
            DY[e] = (gi < n0 - GW && gj < n1 - GW && gk < n2)                                                 \
                        ? d1(&cur[(xx + GW) + HX * ((yy + GW) + HY * zz)], HX, gj, n1, lo1, hi1, a1)          \
                        : 0.0;                                                                                 \
        }                                                                                                      \
        __syncthreads();                                                                                       \
    }

#define BSSN_STAGE_END(F_, S_, FN_)                                                                           \
    if ((FN_) >= 0) {                                                                                          \
        _Pragma("unroll") for (int q = 0; q < NPF; q++) if (t + NT * q < NH) U[((S_) + 1) & 1][t + NT * q] =  \
            pf_##S_[q];                                                                                        \
    }                                                                                                          \
    __syncthreads();

#define BSSN_D1(F_, D_)                                                                                        \
    ((R)((D_) == 0 ? d1(&cur[c], 1, i, n0, lo0, hi0, a0)                                                      \
                   : (D_) == 1 ? d1(&cur[c], HX, j, n1, lo1, hi1, a1) : d1(&cur[c], HX * HY, k, n2, lo2, hi2, a2)))

#define BSSN_ADV(F_)                                                                                           \
    ((R)fused_adv(&cur[c], i, j, k, n0, n1, n2, lo0, hi0, lo1, hi1, lo2, hi2, a0, a1, a2, bv0, bv1, bv2, UPWIND))

#define FUSED_D2(S_)                                                                                           \
    ((S_) == 0   ? d2(&cur[c], 1, i, n0, lo0, hi0, a0)                                                           \
     : (S_) == 3 ? d2(&cur[c], HX, j, n1, lo1, hi1, a1)                                                          \
     : (S_) == 5 ? d2(&cur[c], HX * HY, k, n2, lo2, hi2, a2)                                                     \
     : (S_) == 1 ? d1(&DX[lx + TX * ((ly + GW) + HY * (lz + GW))], TX, j, n1, lo1, hi1, a1)                      \
     : (S_) == 2 ? d1(&DX[lx + TX * ((ly + GW) + HY * (lz + GW))], TX * HY, k, n2, lo2, hi2, a2)                 \
                 : d1(&DY[lx + TX * (ly + TY * (lz + GW))], TX * TY, k, n2, lo2, hi2, a2))

#define BSSN_D2(F_, S_) ((R)FUSED_D2(S_))

    // igt^{lm} d_l d_m of the current (conformal-metric) field, igt from the
    // point's gt values
    auto fused_lap = [&]() -> double {
        double g_[6], ig_[6], h_[6];
#pragma unroll
        for (int q_ = 0; q_ < 6; q_++) g_[q_] = valid ? in.p[12 + q_][pp] : (q_ == 0 || q_ == 3 || q_ == 5 ? 1.0 : 0.0);
        h_[0] = FUSED_D2(0);
        h_[1] = FUSED_D2(1);
        h_[2] = FUSED_D2(2);
        h_[3] = FUSED_D2(3);
        h_[4] = FUSED_D2(4);
        h_[5] = FUSED_D2(5);
        inverse_metric(g_, ig_);
        return metric_contract(ig_, h_);
    };
#define BSSN_LAP(F_) ((R)fused_lap())

    // Kreiss-Oliger sum (KO_x + KO_y) + KO_z and, on a flagged face, the
    // Sommerfeld value; both in double, from the field's LDS tile.
#define BSSN_FIELD_DONE(F_)                                                                                    \
    double ko_##F_, bc_##F_ = 0.0;                                                                             \
    {                                                                                                          \
        SGR_EXACT                                                                                              \
        const double kx_ = ko(&cur[c], 1, i, n0, lo0, hi0, a0);                                                \
        const double ky_ = ko(&cur[c], HX, j, n1, lo1, hi1, a1);                                               \
        const double kz_ = ko(&cur[c], HX * HY, k, n2, lo2, hi2, a2);                                          \
        ko_##F_ = kx_ + ky_ + kz_;                                                                             \
        if (face)                                                                                              \
            bc_##F_ = sommerfeld_f(x, y, z, inv_r, d1(&cur[c], 1, i, n0, lo0, hi0, a0),                        \
                                   d1(&cur[c], HX, j, n1, lo1, hi1, a1),                                       \
                                   d1(&cur[c], HX * HY, k, n2, lo2, hi2, a2), cur[c], bc_falloff(F_),          \
                                   bc_asym(F_));                                                               \
    }

#define BSSN_OUT(F_, V_)                                                                                       \
    if (valid) out.p[F_][pp] = ko_add_f(face ? bc_##F_ : (double)(V_), sigma, ko_##F_);

#include "gen/bssn_staged.inc"

#undef BSSN_U
#undef BSSN_FIRST_FIELD
#undef BSSN_STAGE_BEGIN
#undef BSSN_STAGE_END
#undef BSSN_D1
#undef BSSN_ADV
#undef BSSN_D2
#undef FUSED_D2
#undef BSSN_LAP
#undef BSSN_FIELD_DONE
#undef BSSN_OUT
}

template <typename R, bool UPWIND>
static hipError_t launch_fused_t(const LaunchArgs &a, hipStream_t s)
{
    hipExtLaunchKernelGGL((k_fused<R, UPWIND>), dim3(a.ntiles), dim3(fused::NT), 0, s, a.ev[0], a.ev[1], 0, a.in,
                          a.out, a.blocks, a.tile_block, a.params);
    return hipGetLastError();
}

hipError_t launch_fused(const LaunchArgs &a, hipStream_t s)
{
    if (a.ntiles == 0) return hipSuccess;
    if (a.mixed) return a.upwind ? launch_fused_t<float, true>(a, s) : launch_fused_t<float, false>(a, s);
    return a.upwind ? launch_fused_t<double, true>(a, s) : launch_fused_t<double, false>(a, s);
}

}  // namespace sgr
