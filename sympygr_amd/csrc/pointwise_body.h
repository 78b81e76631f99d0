// Body of one output-group pointwise kernel; included once per group with
// SGR_PW_NAME (kernel name) and SGR_PW_ALG (generated algebra file) defined.
// One thread per interior point: reads its fields and staged derivatives,
// runs the group's part of the generated algebra in R (double or float),
// applies the Sommerfeld condition on flagged faces and adds sigma * KO in
// double.
template <typename R, typename W, bool UPWIND, bool RK>
__global__ __launch_bounds__(256, 2) void SGR_PW_NAME(FieldPtrs in, OutPtrs out, const BlockDesc *__restrict__ blocks,
                                                   const uint32_t *__restrict__ tile_block,
                                                   const W *__restrict__ ws, const double *__restrict__ wsK,
                                                   uint64_t SS, Params P, RkArgs rk, uint32_t tile0)
{
    const uint32_t wg = xcd_remap(blockIdx.x, gridDim.x);  // XCD-contiguous tiles
    const uint32_t tile = tile0 + (wg >> 1);
    const int l = (int)((wg & 1) * 256 + threadIdx.x);
    const BlockDesc &B = blocks[tile_block[tile]];
    const TileCoord tc = tile_coord(B, tile);
    const int lx = l & 7, ly = (l >> 3) & 7, lz = l >> 6;
    const int i = tc.i0 + lx, j = tc.j0 + ly, k = tc.k0 + lz;
    const int n0 = B.n[0], n1 = B.n[1], n2 = B.n[2];
    if (i >= n0 - GW || j >= n1 - GW || k >= n2 - GW) return;
    // this tile's workspace (layout: ws_tile_base, ws_slot_stride)
    const W *__restrict__ wsb = ws + ws_tile_base(wg >> 1, NSLOT) + l;
    const double *__restrict__ kb = wsK + ws_tile_base(wg >> 1, NF) + l;
    SS = ws_slot_stride(SS);
    const uint64_t pp = B.offset + i + (uint64_t)n0 * (j + (uint64_t)n1 * k);

    // coordinates and eta exactly as bssn/src/rhs.cpp:206-219
    const double x = coord(B.pmin[0], i, B.ax[0].h);
    const double y = coord(B.pmin[1], j, B.ax[1].h);
    const double z = coord(B.pmin[2], k, B.ax[2].h);
    double eta_d = P.eta_const;
    if (P.eta_func) {
        // GR/bssn.py:52: R0 sqrt(igt^ij d_i chi d_j chi) / (1 - chi^p0)^p1
        double g[6], ig[6], h[6];
#pragma unroll
        for (int q = 0; q < 6; q++) g[q] = in.p[12 + q][pp];
        inverse_metric(g, ig);
        const double c0 = (double)ws_get(&wsb[slot_d1(1, 0) * SS]);
        const double c1 = (double)ws_get(&wsb[slot_d1(1, 1) * SS]);
        const double c2 = (double)ws_get(&wsb[slot_d1(1, 2) * SS]);
        h[0] = c0 * c0;
        h[1] = c0 * c1;
        h[2] = c0 * c2;
        h[3] = c1 * c1;
        h[4] = c1 * c2;
        h[5] = c2 * c2;
        eta_d = P.eta_func_r0 * sqrt(metric_contract(ig, h)) *
                pow(1.0 - pow(in.p[1][pp], P.eta_power[0]), -P.eta_power[1]);
    } else {
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
    const R xi1 = (R)P.xi[1], xi2 = (R)P.xi[2];
    const bool rochester = P.rochester;
    const double sigma = P.ko_sigma;

#define BSSN_U(F_) ((R)in.p[F_][pp])
#define BSSN_D1(F_, D_) ((R)ws_get(&wsb[slot_d1(F_, D_) * SS]))
#define BSSN_D2(F_, S_) ((R)ws_get(&wsb[slot_d2(F_, S_) * SS]))
#define BSSN_ADV(F_) ((R)ws_get(&wsb[slot_adv(F_) * SS]))
#define BSSN_LAP(F_) ((R)ws_get(&wsb[slot_lap(F_) * SS]))
#define BSSN_GDB(J_) ((R)ws_get(&wsb[slot_gdb(J_) * SS]))
// boundary value on flagged faces, + sigma KO, then the RHS or RK-stage store
#define SGR_PW_STORE(F_, V_)                                                                                  \
    do {                                                                                                    \
        double val_ = (double)(V_);                                                                         \
        if (face)                                                                                           \
            val_ = sommerfeld(x, y, z, inv_r, (double)ws_get(&wsb[slot_d1(F_, 0) * SS]),                    \
                              (double)ws_get(&wsb[slot_d1(F_, 1) * SS]),                                    \
                              (double)ws_get(&wsb[slot_d1(F_, 2) * SS]),                                    \
                              in.p[F_][pp], bc_falloff(F_), bc_asym(F_));                                   \
        val_ = ko_add(val_, sigma, ws_get(&kb[(F_) * SS]));                                                 \
        if (RK)                                                                                             \
            rk_store(rk, F_, pp, val_);                                                                     \
        else                                                                                                \
            out_put(&out.p[F_][pp], val_);                                                                  \
    } while (0)
// the Rochester gauge replaces the shift (6..8) and B (9..11) equations
#define BSSN_OUT(F_, V_)                                                                                      \
    do {                                                                                                    \
        if (!(rochester && (F_) >= 6 && (F_) <= 11)) SGR_PW_STORE(F_, V_);                                  \
    } while (0)
#define BSSN_ALT_ON rochester
#define BSSN_OUT_ALT(F_, V_) SGR_PW_STORE(F_, V_)
#include SGR_PW_ALG
#undef BSSN_BARRIER
#undef BSSN_U
#undef BSSN_D1
#undef BSSN_D2
#undef BSSN_ADV
#undef BSSN_LAP
#undef BSSN_GDB
#undef BSSN_OUT
#undef BSSN_OUT_ALT
#undef BSSN_ALT_ON
#undef SGR_PW_STORE
}

