// Body of the pointwise kernel; included once per system with SGR_PW_NFLD
// (field count), SGR_PW_NAME (kernel name) and SGR_PW_ALG (generated algebra
// file) defined.  The pointwise algebra at every interior point its tile
// owns.  One
// thread per point: reads the point's field values, the derivatives the
// stencil pass staged and the per-output tails (advection plus sigma times
// Kreiss-Oliger, double), runs the algebra in R (double, or float in mixed
// precision), adds the tail in double and stores the RHS, or the fused RK4
// stage update.  On a flagged-face point the tail is the whole output (the
// Sommerfeld value plus sigma KO) and the algebra's value is discarded.
template <typename R, typename W, bool RK>
__global__ __launch_bounds__(256, 2) void SGR_PW_NAME(FieldPtrs in, OutPtrs out, const BlockDesc *__restrict__ blocks,
                                                   const uint32_t *__restrict__ tile_block,
                                                   const W *__restrict__ ws, const double *__restrict__ wsT,
                                                   Params P, RkArgs rk, uint32_t tile0)
{
    const uint32_t wg = xcd_remap(blockIdx.x, gridDim.x);  // XCD-contiguous tiles
    const uint32_t tile = tile0 + (wg >> 1);
    const int l = (int)((wg & 1) * 256 + threadIdx.x);
    const BlockDesc &B = blocks[tile_block[tile]];
    const TileCoord tc = tile_coord(B, tile);
    // point l of the tile, x fastest (the stencil pass's order)
    const int TX = B.tdim[0], TY = B.tdim[1];
    if (l >= TX * TY * B.tdim[2]) return;
    const int r = div_q(l, make_div(TX)), lz = div_q(r, make_div(TY));
    const int lx = l - r * TX, ly = r - lz * TY;
    const int i = tc.i0 + lx, j = tc.j0 + ly, k = tc.k0 + lz;
    const int n0 = B.n[0], n1 = B.n[1], n2 = B.n[2];
    if (i >= n0 - GW || j >= n1 - GW || k >= n2 - GW) return;
    if (i < tc.oi || j < tc.oj || k < tc.ok) return;  // owned by the previous tile (tile_coord)
    const uint32_t bf = B.bflag;
    const bool face = bf && (((bf & 1u) && i == GW) || ((bf & 2u) && i == n0 - GW - 1) ||
                             ((bf & 4u) && j == GW) || ((bf & 8u) && j == n1 - GW - 1) ||
                             ((bf & 16u) && k == GW) || ((bf & 32u) && k == n2 - GW - 1));
    // this tile's workspace (layout: ws_tile_base, slot stride WS_SS)
    const W *__restrict__ wsb = ws + ws_tile_base(wg >> 1, rs_nslot(SGR_PW_NFLD)) + l;
    const double *__restrict__ tb = wsT + ws_tile_base(wg >> 1, SGR_PW_NFLD) + l;
    constexpr uint64_t SS = WS_SS;
    const uint64_t pp = B.offset + i + (uint64_t)n0 * (j + (uint64_t)n1 * k);

    double eta_d;
    if (P.eta_func) {
        double dchi[3], g[6];
#pragma unroll
        for (int d = 0; d < 3; d++) dchi[d] = (double)ws_get(&wsb[rs_d1(1, d) * SS]);
#pragma unroll
        for (int q = 0; q < 6; q++) g[q] = in.p[12 + q][pp];
        eta_d = eta_func_at(P, g, dchi, in.p[1][pp]);
    } else {
        eta_d = eta_const_at(P, coord(B.pmin[0], i, B.ax[0].h), coord(B.pmin[1], j, B.ax[1].h),
                             coord(B.pmin[2], k, B.ax[2].h));
    }
    const R lambda0 = (R)P.lambda[0], lambda1 = (R)P.lambda[1], lambda2 = (R)P.lambda[2], lambda3 = (R)P.lambda[3];
    const R lambda_f0 = (R)P.lambda_f[0], lambda_f1 = (R)P.lambda_f[1];
    const R eta = (R)eta_d;
    const R xi1 = (R)P.xi[1], xi2 = (R)P.xi[2];
    const R kappa1 = (R)P.kappa[0], kappa2 = (R)P.kappa[1], kappa3 = (R)P.kappa[2];
    (void)lambda0, (void)lambda1, (void)lambda2, (void)lambda3, (void)lambda_f0, (void)lambda_f1, (void)xi1, (void)xi2;
    (void)kappa1, (void)kappa2, (void)kappa3;
    const bool rochester = P.rochester;

#define BSSN_U(F_) ((R)in.p[F_][pp])
#define BSSN_D1(F_, D_) ((R)ws_get(&wsb[rs_d1(F_, D_) * SS]))
#define BSSN_D2(F_, S_) ((R)ws_get(&wsb[rs_d2(F_, S_) * SS]))
#define BSSN_LAP(F_) ((R)ws_get(&wsb[rs_lap(F_) * SS]))
#define BSSN_GDB(J_) ((R)ws_get(&wsb[rs_gdb(J_) * SS]))
// algebra value + tail (double), then the RHS or RK-stage store
#define SGR_PW_STORE(F_, V_)                                                                                  \
    do {                                                                                                    \
        const double tl_ = ws_get(&tb[rs_tail(F_) * SS]);                                                   \
        const double val_ = face ? tl_ : tail_add((double)(V_), tl_);                                        \
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
#undef BSSN_LAP
#undef BSSN_GDB
#undef BSSN_OUT
#undef BSSN_OUT_ALT
#undef BSSN_ALT_ON
#undef SGR_PW_STORE
}
