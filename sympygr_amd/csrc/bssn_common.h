// Shared host/device definitions for the BSSN RHS kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sgr {

constexpr int NF = 24;     // evolved BSSN fields, order of bssn/include/grDef.h:31
constexpr int NF_CCZ4 = 25;  // CCZ4: the BSSN fields (Gt holding Gamma-hat) and Theta last
constexpr int NFMAX = NF_CCZ4;
constexpr int GW = 3;      // ghost width (bssn/include/block.h:21)
constexpr int T = 8;       // tile edge (interior points)
constexpr int TH = T + 2 * GW;  // tile edge with halo
constexpr int TILE_PTS = T * T * T;

// Fields with second derivatives (bssn/scripts/bssnrhs_derivs.h):
// alpha, beta0..2, chi, gt0..5.  D2Q[f] = position in that list or -1.
constexpr int ND2F = 11;
__host__ __device__ constexpr int d2q(int f)
{
    return f == 0 ? 0 : f == 6 ? 1 : f == 7 ? 2 : f == 8 ? 3 : f == 1 ? 4 : (f >= 12 && f <= 17) ? 5 + (f - 12) : -1;
}
__host__ __device__ constexpr int d2field(int q)
{
    return q == 0 ? 0 : q == 1 ? 6 : q == 2 ? 7 : q == 3 ? 8 : q == 4 ? 1 : 12 + (q - 5);
}

// Derivative workspace slots, one plane of NP points each:
//   [0,72)     first derivative of field f along d: f*3+d
//   [72,84)    second derivatives of alpha and chi: 72 + 6*(f == chi) + s
//              (s: xx xy xz yy yz zz)
//   [84,93)    igt^{lm} d_l d_m f for the shift (84 + f - 6) and the
//              conformal metric (87 + f - 12), igt the inverse conformal
//              metric of the point
//   [93,96)    d_j (d_k beta^k), 93 + j
//   [96,120)   advection of field f, 96 + f:
//              (beta^0 ad_0 f + beta^1 ad_1 f) + beta^2 ad_2 f, with ad the
//              upwinded stencil (upwind mode) or the centred one
constexpr int SLOT_D1 = 0;
constexpr int SLOT_D2 = 72;
constexpr int SLOT_LAP = 84;
constexpr int SLOT_GDB = 93;
constexpr int SLOT_ADV = 96;
constexpr int NSLOT = 120;

__host__ __device__ constexpr int slot_d1(int f, int d) { return SLOT_D1 + f * 3 + d; }
__host__ __device__ constexpr int slot_d2(int f, int s) { return SLOT_D2 + (f == 1 ? 6 : 0) + s; }
__host__ __device__ constexpr int slot_lap(int f) { return f <= 8 ? SLOT_LAP + (f - 6) : SLOT_LAP + 3 + (f - 12); }
__host__ __device__ constexpr int slot_gdb(int j) { return SLOT_GDB + j; }
__host__ __device__ constexpr int slot_adv(int f) { return SLOT_ADV + f; }
// symmetric slot of (j, k): xx xy xz yy yz zz
__host__ __device__ constexpr int sym_slot(int j, int k)
{
    return j > k ? sym_slot(k, j) : (j == 0 ? k : (j == 1 ? 2 + k : 5));
}

// Derivative workspace of the RHS pass (k_stencil -> k_pointwise).  The
// stencil pass stages what the pointwise algebra reads at every interior
// point that is not on a flagged face (it stores the outputs of face points
// itself):
//   [0,45)    first derivative of field f along d, 3 q(f) + d, for
//             alpha chi K Gt0..2 beta0..2 (q = f) and gt0..5 (q = f - 3)
//             (B and At enter the algebra without derivatives)
//   [45,57)   second derivatives of alpha (45 + s) and chi (51 + s)
//   [57,66)   igt^{lm} d_l d_m f of the shift (57 + f - 6) and gt (60 + f - 12)
//   [66,69)   d_j (d_k beta^k), 66 + j
// and, always in double, the tail of each of the 24 outputs: the terms
// linear in the field's own advective and Kreiss-Oliger stencils, added
// after the algebra, c_f beta.ad f + sigma KO f, with c_f the advection's
// coefficient in the equation (lambda0 for alpha, lambda1 or, in the
// Rochester gauge, xi1 for the shift, 1 otherwise) and for B^i
// (1 - lambda3) beta.ad Gt^i + lambda2 beta.ad B^i + sigma KO B^i
// (sigma KO B^i alone in the Rochester gauge, where B is not evolved).
// CCZ4 adds the first derivatives of Theta (slots 69..71) and its tail.
constexpr int RS_NSLOT = 69;
__host__ __device__ constexpr int rs_nslot(int nfld) { return nfld == NF_CCZ4 ? RS_NSLOT + 3 : RS_NSLOT; }
__host__ __device__ constexpr bool rs_has_d1(int f) { return f <= 8 || (f >= 12 && f <= 17) || f == 24; }
__host__ __device__ constexpr int rs_d1(int f, int d) { return f == 24 ? RS_NSLOT + d : 3 * (f <= 8 ? f : f - 3) + d; }
__host__ __device__ constexpr int rs_d2(int f, int s) { return (f == 1 ? 51 : 45) + s; }
__host__ __device__ constexpr int rs_lap(int f) { return f <= 8 ? 57 + (f - 6) : 60 + (f - 12); }
__host__ __device__ constexpr int rs_gdb(int j) { return 66 + j; }
__host__ __device__ constexpr int rs_tail(int f) { return f; }

// Workspace layout, tile-major: slot s of point p (x fastest in the 8^3
// tile) of tile t at ws[t * TILE_PTS * nslots + s * TILE_PTS + p], so a
// tile's slots are one contiguous run and every wave reads and writes 512
// contiguous bytes per slot.
__host__ __device__ constexpr uint64_t ws_tile_base(uint64_t tile, int nslots)
{
    return tile * (uint64_t)TILE_PTS * (uint64_t)nslots;
}
constexpr uint64_t WS_SS = TILE_PTS;  // slot stride

// The derivative workspace is streamed: written once by the derivative pass,
// read once per pointwise kernel, far larger than L2.  Non-temporal accesses
// keep it from displacing the field tiles the derivative pass re-reads.
#ifndef SGR_WS_NONTEMPORAL
#define SGR_WS_NONTEMPORAL 1
#endif
template <typename X>
__device__ __forceinline__ void ws_put(X *p, X v)
{
#if SGR_WS_NONTEMPORAL
    __builtin_nontemporal_store(v, p);
#else
    *p = v;
#endif
}
template <typename X>
__device__ __forceinline__ X ws_get(const X *p)
{
#if SGR_WS_NONTEMPORAL
    return __builtin_nontemporal_load(p);
#else
    return *p;
#endif
}

// RHS and RK-stage outputs: plain stores, or non-temporal ones.
#ifndef SGR_OUT_NONTEMPORAL
#define SGR_OUT_NONTEMPORAL 0
#endif
__device__ __forceinline__ void out_put(double *p, double v)
{
#if SGR_OUT_NONTEMPORAL
    __builtin_nontemporal_store(v, p);
#else
    *p = v;
#endif
}

// Inverse conformal metric (adjugate / determinant) from gt0..5 (xx xy xz yy
// yz zz), the same expressions as the pointwise algebra's igt; and the
// contraction sum_s w_s igt_s h_s of a symmetric second-derivative set h with
// w = (1,2,2,1,2,1), summed in slot order.
__host__ __device__ inline void inverse_metric(const double g[6], double ig[6])
{
#pragma clang fp contract(off)
    const double c0 = g[3] * g[5] - g[4] * g[4];
    const double c1 = g[2] * g[4] - g[1] * g[5];
    const double c2 = g[1] * g[4] - g[2] * g[3];
    const double c3 = g[0] * g[5] - g[2] * g[2];
    const double c4 = g[1] * g[2] - g[0] * g[4];
    const double c5 = g[0] * g[3] - g[1] * g[1];
    const double idet = 1.0 / (g[0] * c0 + g[1] * c1 + g[2] * c2);
    ig[0] = c0 * idet;
    ig[1] = c1 * idet;
    ig[2] = c2 * idet;
    ig[3] = c3 * idet;
    ig[4] = c4 * idet;
    ig[5] = c5 * idet;
}
__host__ __device__ inline double metric_contract(const double ig[6], const double h[6])
{
#pragma clang fp contract(off)
    return ((((ig[0] * h[0] + 2.0 * (ig[1] * h[1])) + 2.0 * (ig[2] * h[2])) + ig[3] * h[3]) + 2.0 * (ig[4] * h[4])) +
           ig[5] * h[5];
}

// Sommerfeld falloff / asymptotic value per field (bssn/src/rhs.cpp:261-320);
// the CCZ4 Theta (24) falls off like K.
__host__ __device__ constexpr double bc_falloff(int f)
{
    return (f >= 3 && f <= 5) || (f >= 18 && f <= 23) ? 2.0 : 1.0;
}
__host__ __device__ constexpr double bc_asym(int f)
{
    return (f == 0 || f == 1 || f == 12 || f == 15 || f == 17) ? 1.0 : 0.0;
}

// Per-axis stencil constants, computed on the host with the reference's
// expressions (bssn/src/derivs.cpp) so the device sees identical doubles.
struct AxisConst {
    double h;        // (pmax - pmin)/(n - 1)            rhs.cpp:144-146
    double by60;     // (1/h)/60                         derivs.cpp:2061
    double by12;     // (1/h)/12                         derivs.cpp:2060
    double sq180;    // (1/(h h))/180                    derivs.cpp:2448
    double sq12;     // (1/(h h))/12                     derivs.cpp:2449
    double by2;      // 0.5 (1/h)                        derivs.cpp:501
    double ko_pre;   // -1/64/h                          derivs.cpp:919
    double smr3, smr2, smr1;  // 59/48*64*h, 43/48*64*h, 49/48*64*h   derivs.cpp:921-923
};

// One padded block as the kernels see it.  The RHS pass tiles a block with
// tiles of tdim[0] x tdim[1] x tdim[2] interior points (at most TILE_PTS;
// 8^3, or the shape tuned for the block's size); the constraint pass always
// uses 8^3 tiles.
struct BlockDesc {
    uint64_t offset;     // start of the block in every field array
    uint64_t ws_base;    // first workspace point of the block (tile-major)
    int32_t n[3];        // padded sizes
    int32_t ntile[3];    // RHS tiles per axis = ceil((n-6)/tdim)
    int32_t tdim[3];     // RHS tile extents
    int32_t ntile8[3];   // constraint-pass tiles per axis = ceil((n-6)/8)
    uint32_t bflag;
    uint32_t tile_base;  // first RHS tile id of this block
    uint32_t tile_base8; // first constraint-pass tile id of this block
    double pmin[3];
    AxisConst ax[3];
};

// Limits of a general RHS tile shape (k_stencil_gen's LDS arrays): halo tile
// (tx+6)(ty+6)(tz+6), D_x stage tx(ty+6)(tz+6), D_y stage tx ty (tz+6).
constexpr int GEN_U_CAP = 3456;
constexpr int GEN_DX_CAP = 3200;
constexpr int GEN_DY_CAP = 1600;
__host__ __device__ constexpr bool tile_shape_ok(int tx, int ty, int tz)
{
    return tx >= 1 && ty >= 1 && tz >= 1 && tx * ty * tz <= TILE_PTS &&
           (tx + 6) * (ty + 6) * (tz + 6) <= GEN_U_CAP && tx * (ty + 6) * (tz + 6) <= GEN_DX_CAP &&
           tx * ty * (tz + 6) <= GEN_DY_CAP;
}

// Physics parameters (bssn/src/parameters.cpp:99-109).
struct Params {
    double lambda[4];
    double lambda_f[2];
    double eta_const, eta_r0, eta_damping_exp;
    double ko_sigma;
    int rochester;       // Rochester puncture gauge (shift equation, B not evolved)
    int eta_func;        // eta from the eta function instead of the constant
    double xi[3];        // BSSN_XI
    double eta_func_r0;  // BSSN_ETA_R0
    double eta_power[2]; // BSSN_ETA_POWER
    double kappa[3];     // CCZ4 damping kappa1..3
};

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own
// L2.  Remap the launch index so each XCD works on one contiguous range of
// tiles: x-neighbouring tiles then share cache lines and halo rows in the same
// L2 instead of fetching them from HBM once per XCD.  Bijective for any n.
__device__ __forceinline__ uint32_t xcd_remap(uint32_t b, uint32_t n)
{
    constexpr uint32_t NX = 8;
    const uint32_t q = n / NX, r = n % NX;
    const uint32_t x = b % NX, i = b / NX;
    return x * q + (x < r ? x : r) + i;
}

struct FieldPtrs {
    const double *p[NFMAX];
};
struct OutPtrs {
    double *p[NFMAX];
};

}  // namespace sgr
