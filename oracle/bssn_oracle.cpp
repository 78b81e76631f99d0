// Test infrastructure only -- CPU restatement of the reference BSSN RHS.
//
// It mirrors the reference's structure on purpose, because it is both the
// parity checker and the timed CPU baseline: per block, full-array stencil
// sweeps into temporaries (bssn/scripts/bssnrhs_derivs.h,
// bssnrhs_derivs_adv.h), then the pointwise loop (bssn/src/rhs.cpp:206-258),
// then Sommerfeld boundary conditions on flagged faces (rhs.cpp:261-320,
// 833-977), then Kreiss-Oliger sweeps and the dissipation add
// (rhs.cpp:324-369).  The pointwise algebra is the project's own SymPy
// statement (gen/bssn_algebra.inc); everything else is written here.
//
// Stencil taps are applied left to right in the order the reference writes
// them, with contraction off (-ffp-contract=off), so each sweep is bit-equal
// to the reference sweep.  The tap tables have to be the reference's: they
// are the finite-difference operators being restated.
#include "oracle.h"

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

struct Tap {
    int off;   // offset from the point along the sweep axis
    double c;  // coefficient
};

struct Stencil {
    int n;
    Tap t[7];
};

double apply(const Stencil &s, const double *u, std::ptrdiff_t p, std::ptrdiff_t stride)
{
    double acc = s.t[0].c * u[p + s.t[0].off * stride];
    for (int m = 1; m < s.n; m++) acc = acc + s.t[m].c * u[p + s.t[m].off * stride];
    return acc;
}

// ---- 6th-order centred first derivative and its one-sided face closures,
// bssn/src/derivs.cpp:2055-2434 (deriv64_{x,y,z}).
const Stencil D1_INT = {6, {{-3, -1.0}, {-2, 9.0}, {-1, -45.0}, {1, 45.0}, {2, -9.0}, {3, 1.0}}};
const Stencil D1_L[3] = {
    {5, {{0, -25.0}, {1, 48.0}, {2, -36.0}, {3, 16.0}, {4, -3.0}}},
    {5, {{-1, -3.0}, {0, -10.0}, {1, 18.0}, {2, -6.0}, {3, 1.0}}},
    {4, {{-2, 1.0}, {-1, -8.0}, {1, 8.0}, {2, -1.0}}},
};
const Stencil D1_R[3] = {  // at ie-3, ie-2, ie-1 (ie = n-3)
    {4, {{-2, 1.0}, {-1, -8.0}, {1, 8.0}, {2, -1.0}}},
    {5, {{-3, -1.0}, {-2, 6.0}, {-1, -18.0}, {0, 10.0}, {1, 3.0}}},
    {5, {{-4, 3.0}, {-3, -16.0}, {-2, 36.0}, {-1, -48.0}, {0, 25.0}}},
};

// ---- 6th-order centred second derivative and closures,
// bssn/src/derivs.cpp:2443-2833 (deriv64_{xx,yy,zz}).
const Stencil D2_INT = {7, {{-3, 2.0}, {-2, -27.0}, {-1, 270.0}, {0, -490.0}, {1, 270.0}, {2, -27.0}, {3, 2.0}}};
const Stencil D2_L[3] = {
    {5, {{0, 35.0}, {1, -104.0}, {2, 114.0}, {3, -56.0}, {4, 11.0}}},
    {5, {{-1, 11.0}, {0, -20.0}, {1, 6.0}, {2, 4.0}, {3, -1.0}}},
    {5, {{-2, -1.0}, {-1, 16.0}, {0, -30.0}, {1, 16.0}, {2, -1.0}}},
};
const Stencil D2_R[3] = {
    {5, {{-2, -1.0}, {-1, 16.0}, {0, -30.0}, {1, 16.0}, {2, -1.0}}},
    {5, {{-3, -1.0}, {-2, 4.0}, {-1, 6.0}, {0, -20.0}, {1, 11.0}}},
    {5, {{-4, 11.0}, {-3, -56.0}, {-2, 114.0}, {-1, -104.0}, {0, 35.0}}},
};

// ---- 4th-order shift-upwinded advective derivative, bssn/src/derivs.cpp:495-906
// (deriv42adv_{x,y,z}); "P" = beta > 0, "M" = otherwise.  Scale 1/(12h)
// for the 5-point forms and 1/(2h) for the 2/3-point forms.
const Stencil AD_INT_P = {5, {{-1, -3.0}, {0, -10.0}, {1, 18.0}, {2, -6.0}, {3, 1.0}}};
const Stencil AD_INT_M = {5, {{-3, -1.0}, {-2, 6.0}, {-1, -18.0}, {0, 10.0}, {1, 3.0}}};
const Stencil AD_FWD2 = {3, {{0, -3.0}, {1, 4.0}, {2, -1.0}}};  // (-3,4,-1)/2h forward
const Stencil AD_CEN2 = {2, {{-1, -1.0}, {1, 1.0}}};           // (-1,0,1)/2h
const Stencil AD_BWD2 = {3, {{-2, 1.0}, {-1, -4.0}, {0, 3.0}}}; // (1,-4,3)/2h backward

// ---- Kreiss-Oliger, bssn/src/derivs.cpp:915-1319 (ko_deriv42_{x,y,z}).
const Stencil KO_INT = {7, {{-3, -1.0}, {-2, 6.0}, {-1, -15.0}, {0, 20.0}, {1, -15.0}, {2, 6.0}, {3, -1.0}}};
const Stencil KO_L[3] = {  // divided by smr3, smr2, smr1
    {4, {{3, 1.0}, {2, -3.0}, {1, 3.0}, {0, -1.0}}},
    {5, {{3, 1.0}, {2, -6.0}, {1, 12.0}, {0, -10.0}, {-1, 3.0}}},
    {6, {{3, 1.0}, {2, -6.0}, {1, 15.0}, {0, -19.0}, {-1, 12.0}, {-2, -3.0}}},
};
const Stencil KO_R[3] = {  // at ie-3, ie-2, ie-1, divided by spr1, spr2, spr3
    {6, {{-3, 1.0}, {-2, -6.0}, {-1, 15.0}, {0, -19.0}, {1, 12.0}, {2, -3.0}}},
    {5, {{-3, 1.0}, {-2, -6.0}, {-1, 12.0}, {0, -10.0}, {1, 3.0}}},
    {4, {{-3, 1.0}, {-2, -3.0}, {-1, 3.0}, {0, -1.0}}},
};

struct Box {
    int lo[3], hi[3];
};

struct Grid {
    int nx, ny, nz;
    std::ptrdiff_t idx(int i, int j, int k) const { return i + (std::ptrdiff_t)nx * (j + (std::ptrdiff_t)ny * k); }
    std::ptrdiff_t stride(int a) const { return a == 0 ? 1 : (a == 1 ? nx : (std::ptrdiff_t)nx * ny); }
    int n(int a) const { return a == 0 ? nx : (a == 1 ? ny : nz); }
};

Box interior(const Grid &g) { return {{3, 3, 3}, {g.nx - 3, g.ny - 3, g.nz - 3}}; }

// Visit every point of box with the sweep-axis coordinate fixed to `pos`
// (pos < 0: visit the whole box).
template <class F>
void for_box(const Grid &g, const Box &b, int axis, int pos, F f)
{
    Box c = b;
    if (pos >= 0) { c.lo[axis] = pos; c.hi[axis] = pos + 1; }
    for (int k = c.lo[2]; k < c.hi[2]; k++)
        for (int j = c.lo[1]; j < c.hi[1]; j++)
            for (int i = c.lo[0]; i < c.hi[0]; i++) f(g.idx(i, j, k), i, j, k);
}

// deriv64_{x,y,z}: x sweeps i in [3,n-3) and all j,k; y sweeps i,j interior and
// all k; z sweeps the interior (derivs.cpp:2066-2071, 2197-2202, 2324-2329).
// The wider ranges give the mixed derivatives valid inner values in ghost rows.
void first(double *D, const double *u, double h, const Grid &g, unsigned bflag, int axis)
{
    const double ih = 1.0 / h;
    const double by12 = ih / 12.0, by60 = ih / 60.0;
    Box b = interior(g);
    if (axis == 0) { b.lo[1] = 0; b.hi[1] = g.ny; b.lo[2] = 0; b.hi[2] = g.nz; }
    if (axis == 1) { b.lo[2] = 0; b.hi[2] = g.nz; }
    const std::ptrdiff_t s = g.stride(axis);
    for_box(g, b, axis, -1, [&](std::ptrdiff_t p, int, int, int) { D[p] = apply(D1_INT, u, p, s) * by60; });
    const int ie = g.n(axis) - 3;
    if (bflag & (1u << (2 * axis)))
        for (int m = 0; m < 3; m++)
            for_box(g, b, axis, 3 + m, [&](std::ptrdiff_t p, int, int, int) { D[p] = apply(D1_L[m], u, p, s) * by12; });
    if (bflag & (1u << (2 * axis + 1)))
        for (int m = 0; m < 3; m++)
            for_box(g, b, axis, ie - 3 + m, [&](std::ptrdiff_t p, int, int, int) { D[p] = apply(D1_R[m], u, p, s) * by12; });
}

void second(double *D, const double *u, double h, const Grid &g, unsigned bflag, int axis)
{
    const double ih2 = 1.0 / (h * h);
    const double by180 = ih2 / 180.0, by12 = ih2 / 12.0;
    const Box b = interior(g);
    const std::ptrdiff_t s = g.stride(axis);
    for_box(g, b, axis, -1, [&](std::ptrdiff_t p, int, int, int) { D[p] = apply(D2_INT, u, p, s) * by180; });
    const int ie = g.n(axis) - 3;
    if (bflag & (1u << (2 * axis)))
        for (int m = 0; m < 3; m++)
            for_box(g, b, axis, 3 + m, [&](std::ptrdiff_t p, int, int, int) { D[p] = apply(D2_L[m], u, p, s) * by12; });
    if (bflag & (1u << (2 * axis + 1)))
        for (int m = 0; m < 3; m++)
            for_box(g, b, axis, ie - 3 + m, [&](std::ptrdiff_t p, int, int, int) { D[p] = apply(D2_R[m], u, p, s) * by12; });
}

void advective(double *D, const double *u, double h, const Grid &g, const double *beta, unsigned bflag, int axis)
{
    const double ih = 1.0 / h;
    const double by2 = 0.50 * ih, by12 = ih / 12.0;
    const Box b = interior(g);
    const std::ptrdiff_t s = g.stride(axis);
    for_box(g, b, axis, -1, [&](std::ptrdiff_t p, int, int, int) {
        D[p] = (beta[p] > 0.0 ? apply(AD_INT_P, u, p, s) : apply(AD_INT_M, u, p, s)) * by12;
    });
    const int ie = g.n(axis) - 3;
    if (bflag & (1u << (2 * axis))) {
        for_box(g, b, axis, 3, [&](std::ptrdiff_t p, int, int, int) { D[p] = apply(AD_FWD2, u, p, s) * by2; });
        for_box(g, b, axis, 4, [&](std::ptrdiff_t p, int, int, int) {
            D[p] = (beta[p] > 0.0 ? apply(AD_FWD2, u, p, s) : apply(AD_CEN2, u, p, s)) * by2;
        });
        for_box(g, b, axis, 5, [&](std::ptrdiff_t p, int, int, int) {
            D[p] = beta[p] > 0.0 ? apply(AD_INT_P, u, p, s) * by12 : apply(AD_BWD2, u, p, s) * by2;
        });
    }
    if (bflag & (1u << (2 * axis + 1))) {
        for_box(g, b, axis, ie - 3, [&](std::ptrdiff_t p, int, int, int) {
            D[p] = beta[p] > 0.0 ? apply(AD_FWD2, u, p, s) * by2 : apply(AD_INT_M, u, p, s) * by12;
        });
        for_box(g, b, axis, ie - 2, [&](std::ptrdiff_t p, int, int, int) {
            D[p] = (beta[p] > 0.0 ? apply(AD_CEN2, u, p, s) : apply(AD_BWD2, u, p, s)) * by2;
        });
        for_box(g, b, axis, ie - 1, [&](std::ptrdiff_t p, int, int, int) { D[p] = apply(AD_BWD2, u, p, s) * by2; });
    }
}

void kreiss_oliger(double *D, const double *u, double h, const Grid &g, unsigned bflag, int axis)
{
    const double pre = -1.0 / 64.0 / h;
    const double smr3 = 59.0 / 48.0 * 64 * h;
    const double smr2 = 43.0 / 48.0 * 64 * h;
    const double smr1 = 49.0 / 48.0 * 64 * h;
    const double sl[3] = {smr3, smr2, smr1};
    const double sr[3] = {smr1, smr2, smr3};
    const Box b = interior(g);
    const std::ptrdiff_t s = g.stride(axis);
    for_box(g, b, axis, -1, [&](std::ptrdiff_t p, int, int, int) { D[p] = pre * apply(KO_INT, u, p, s); });
    const int ie = g.n(axis) - 3;
    if (bflag & (1u << (2 * axis)))
        for (int m = 0; m < 3; m++)
            for_box(g, b, axis, 3 + m, [&](std::ptrdiff_t p, int, int, int) { D[p] = apply(KO_L[m], u, p, s) / sl[m]; });
    if (bflag & (1u << (2 * axis + 1)))
        for (int m = 0; m < 3; m++)
            for_box(g, b, axis, ie - 3 + m, [&](std::ptrdiff_t p, int, int, int) { D[p] = apply(KO_R[m], u, p, s) / sr[m]; });
}

// ---- TVB dissipation, bssn/src/derivs.cpp:1348-2054 (disstvb3_{x,y,z},
// disstvb5_{x,y,z}), weighted by the characteristic speeds of
// bssn/src/rhs.cpp:984-1028 (max_spacetime_speeds).  Each tap of an interior
// form is multiplied by the speed along the sweep axis at that tap (TVB3), or
// each difference u[o+1] - u[o] by the larger of the two speeds (TVB5).
// Faces flagged as domain boundaries take the Kreiss-Oliger one-sided
// closures (KO_L / KO_R: the reference writes the same formulas).  Where the
// reference's result is undefined this restatement makes a choice:
//   - the speeds are evaluated at every node, ghost nodes included (the
//     reference fills interior nodes only, and its stencils read ghost ones);
//   - TVB3 at the first and last interior node before an unflagged face uses
//     the interior form (the reference leaves the array's previous contents).
// The reference scales the TVB5 interior form by 2/(75 h) in x and y and by
// -1/(64 h) in z (derivs.cpp:1668, 1796, 1929); kept as it is.
const Stencil TVB3_INT = {5, {{-2, 1.0}, {-1, -4.0}, {0, 6.0}, {1, -4.0}, {2, 1.0}}};
const Stencil TVB5_FIRST = {6, {{3, 1.0}, {2, -6.0}, {1, 15.0}, {0, -19.0}, {-1, 12.0}, {-2, -3.0}}};  // / smr1
const Stencil TVB5_LAST = {6, {{-4, 1.0}, {-3, -6.0}, {-2, 15.0}, {-1, -19.0}, {0, 12.0}, {1, -3.0}}};  // / spr1
const Tap TVB5_DIFF[6] = {{2, 1.0}, {1, -5.0}, {0, 10.0}, {-1, -10.0}, {-2, 5.0}, {-3, -1.0}};

// sum of (c lam) u over the taps, left to right (the reference's order)
double apply_w(const Stencil &s, const double *u, const double *lam, std::ptrdiff_t p, std::ptrdiff_t stride)
{
    std::ptrdiff_t q = p + s.t[0].off * stride;
    double acc = s.t[0].c * lam[q] * u[q];
    for (int m = 1; m < s.n; m++) {
        q = p + s.t[m].off * stride;
        acc = acc + s.t[m].c * lam[q] * u[q];
    }
    return acc;
}

// std::max(a, b)
double larger(double a, double b) { return a < b ? b : a; }

double tvb5_diff(const double *u, const double *lam, std::ptrdiff_t p, std::ptrdiff_t s)
{
    double acc = 0.0;
    for (int m = 0; m < 6; m++) {
        const std::ptrdiff_t q0 = p + TVB5_DIFF[m].off * s, q1 = q0 + s;
        const double t = TVB5_DIFF[m].c * larger(lam[q1], lam[q0]) * (u[q1] - u[q0]);
        acc = m ? acc + t : t;
    }
    return acc;
}

void tvb(double *D, const double *u, const double *lam, double h, const Grid &g, unsigned bflag, int axis, int order)
{
    const double smr3 = 59.0 / 48.0 * 64 * h;
    const double smr2 = 43.0 / 48.0 * 64 * h;
    const double smr1 = 49.0 / 48.0 * 64 * h;
    const double sl[3] = {smr3, smr2, smr1};
    const double sr[3] = {smr1, smr2, smr3};
    const Box b = interior(g);
    const std::ptrdiff_t s = g.stride(axis);
    const int ie = g.n(axis) - 3;
    if (order == 3) {
        const double pre = -1.0 / 12.0 / h;
        for_box(g, b, axis, -1, [&](std::ptrdiff_t p, int, int, int) { D[p] = pre * apply_w(TVB3_INT, u, lam, p, s); });
    } else {
        const double pre = axis == 2 ? -1.0 / 64.0 / h : 2.0 / 75.0 / h;
        for_box(g, b, axis, -1, [&](std::ptrdiff_t p, int, int, int) { D[p] = pre * tvb5_diff(u, lam, p, s); });
        for_box(g, b, axis, 3, [&](std::ptrdiff_t p, int, int, int) { D[p] = apply_w(TVB5_FIRST, u, lam, p, s) / smr1; });
        for_box(g, b, axis, ie - 1, [&](std::ptrdiff_t p, int, int, int) { D[p] = apply_w(TVB5_LAST, u, lam, p, s) / smr1; });
    }
    if (bflag & (1u << (2 * axis)))
        for (int m = 0; m < 3; m++)
            for_box(g, b, axis, 3 + m, [&](std::ptrdiff_t p, int, int, int) { D[p] = apply(KO_L[m], u, p, s) / sl[m]; });
    if (bflag & (1u << (2 * axis + 1)))
        for (int m = 0; m < 3; m++)
            for_box(g, b, axis, ie - 3 + m, [&](std::ptrdiff_t p, int, int, int) { D[p] = apply(KO_R[m], u, p, s) / sr[m]; });
}

// Characteristic speeds at every node of a block (rhs.cpp:984-1028):
// lam_d = max(|-beta^d + t_d|, |-beta^d - t_d|), t_d = alpha sqrt(gtu_dd chi),
// gtu_dd the diagonal cofactors of gt (det gt = 1 assumed), all three set to 1
// where one of them is negative.
void spacetime_speeds(double *const lam[3], const double *const *in, std::size_t n)
{
    for (std::size_t p = 0; p < n; p++) {
        double u11 = in[15][p] * in[17][p] - in[16][p] * in[16][p];
        double u22 = in[12][p] * in[17][p] - in[14][p] * in[14][p];
        double u33 = in[12][p] * in[15][p] - in[13][p] * in[13][p];
        if (u11 < 0.0 || u22 < 0.0 || u33 < 0.0) u11 = u22 = u33 = 1.0;
        const double t[3] = {in[0][p] * std::sqrt(u11 * in[1][p]), in[0][p] * std::sqrt(u22 * in[1][p]),
                             in[0][p] * std::sqrt(u33 * in[1][p])};
        for (int d = 0; d < 3; d++) lam[d][p] = larger(std::fabs(-in[6 + d][p] + t[d]), std::fabs(-in[6 + d][p] - t[d]));
    }
}

// Fields with second derivatives (bssn/scripts/bssnrhs_derivs.h): alpha,
// beta0..2, chi, gt0..5.
const int D2_FIELDS[11] = {0, 6, 7, 8, 1, 12, 13, 14, 15, 16, 17};

// Sommerfeld falloff and asymptotic value per field (bssn/src/rhs.cpp:261-320);
// the CCZ4 Theta (field 24) falls off like K: (1, 0).
const double BC_FALLOFF[25] = {1, 1, 1, 2, 2, 2, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 2, 2, 2, 2, 2, 2, 1};
const double BC_ASYM[25] = {1, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 1, 0, 1, 0, 0, 0, 0, 0, 0, 0};

constexpr int NFMAX = 25;  // BSSN: 24 fields; CCZ4: 25 (Theta last)

struct Workspace {
    std::vector<double> buf;
    double *grad[NFMAX][3];
    double *agrad[NFMAX][3];
    double *grad2[NFMAX][6];
    void resize(std::size_t n, bool upwind, int nf = 24)
    {
        const std::size_t narr = 3 * nf + 66 + (upwind ? 3 * nf : 0);
        if (buf.size() < narr * n) buf.assign(narr * n, 0.0);
        double *p = buf.data();
        for (int f = 0; f < nf; f++)
            for (int d = 0; d < 3; d++) { grad[f][d] = p; p += n; }
        for (int f = 0; f < NFMAX; f++)
            for (int s = 0; s < 6; s++) grad2[f][s] = nullptr;
        for (int q = 0; q < 11; q++)
            for (int s = 0; s < 6; s++) { grad2[D2_FIELDS[q]][s] = p; p += n; }
        for (int f = 0; f < nf; f++)
            for (int d = 0; d < 3; d++) {
                if (upwind) { agrad[f][d] = p; p += n; }
                else agrad[f][d] = grad[f][d];
            }
    }
};

// derivative sweeps (bssnrhs_derivs.h): d_x, d_y, d_z of all fields, pure
// second derivatives and the mixed ones as compositions grad2_0_1 = D_y(grad_0),
// grad2_0_2 = D_z(grad_0), grad2_1_2 = D_z(grad_1); advective derivatives along
// d use beta^d (bssnrhs_derivs_adv.h).
void sweep_all(Workspace &W, const double *const *in, const double *h, const Grid &g, unsigned bflag, bool upwind,
               int nf = 24)
{
    for (int f = 0; f < nf; f++)
        for (int a = 0; a < 3; a++) first(W.grad[f][a], in[f], h[a], g, bflag, a);
    for (int q = 0; q < 11; q++) {
        const int f = D2_FIELDS[q];
        second(W.grad2[f][0], in[f], h[0], g, bflag, 0);
        second(W.grad2[f][3], in[f], h[1], g, bflag, 1);
        second(W.grad2[f][5], in[f], h[2], g, bflag, 2);
        first(W.grad2[f][1], W.grad[f][0], h[1], g, bflag, 1);
        first(W.grad2[f][2], W.grad[f][0], h[2], g, bflag, 2);
        first(W.grad2[f][4], W.grad[f][1], h[2], g, bflag, 2);
    }
    if (upwind)
        for (int f = 0; f < nf; f++)
            for (int a = 0; a < 3; a++) advective(W.agrad[f][a], in[f], h[a], g, in[6 + a], bflag, a);
}

// igt^{lm} d_l d_m of a conformal-metric component at one point: inverse
// metric by adjugate / determinant of gt0..5 (xx xy xz yy yz zz), then
// sum_s w_s igt_s grad2_s, w = (1,2,2,1,2,1), in slot order.  This is the
// first term of the conformal Ricci tensor of GR/dendro.py:495-499 (the
// second derivatives of gt enter the system only in this contraction).
double metric_laplacian(const double *const *in, double *const *grad2, std::ptrdiff_t pp)
{
    const double g0 = in[12][pp], g1 = in[13][pp], g2 = in[14][pp], g3 = in[15][pp], g4 = in[16][pp], g5 = in[17][pp];
    const double c0 = g3 * g5 - g4 * g4;
    const double c1 = g2 * g4 - g1 * g5;
    const double c2 = g1 * g4 - g2 * g3;
    const double c3 = g0 * g5 - g2 * g2;
    const double c4 = g1 * g2 - g0 * g4;
    const double c5 = g0 * g3 - g1 * g1;
    const double idet = 1.0 / (g0 * c0 + g1 * c1 + g2 * c2);
    const double ig[6] = {c0 * idet, c1 * idet, c2 * idet, c3 * idet, c4 * idet, c5 * idet};
    return ((((ig[0] * grad2[0][pp] + 2.0 * (ig[1] * grad2[1][pp])) + 2.0 * (ig[2] * grad2[2][pp])) +
              ig[3] * grad2[3][pp]) + 2.0 * (ig[4] * grad2[4][pp])) + ig[5] * grad2[5][pp];
}

// symmetric second-derivative slot of (j, k): xx xy xz yy yz zz
const int SYM[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};

// igt^ij c_i c_j at one point, igt by adjugate / determinant of gt0..5.
double inverse_metric_quadratic(const double *const *in, std::ptrdiff_t pp, double c0, double c1, double c2)
{
    const double g0 = in[12][pp], g1 = in[13][pp], g2 = in[14][pp], g3 = in[15][pp], g4 = in[16][pp], g5 = in[17][pp];
    const double a0 = g3 * g5 - g4 * g4;
    const double a1 = g2 * g4 - g1 * g5;
    const double a2 = g1 * g4 - g2 * g3;
    const double a3 = g0 * g5 - g2 * g2;
    const double a4 = g1 * g2 - g0 * g4;
    const double a5 = g0 * g3 - g1 * g1;
    const double idet = 1.0 / (g0 * a0 + g1 * a1 + g2 * a2);
    return idet * (a0 * c0 * c0 + 2.0 * a1 * c0 * c1 + 2.0 * a2 * c0 * c2 + a3 * c1 * c1 + 2.0 * a4 * c1 * c2 +
                   a5 * c2 * c2);
}

int rhs_block(double **outv, const double **inv, uint64_t offset, const double *pmin, const double *pmax,
              const uint32_t *sz, uint32_t bflag, const oracle_params *P, Workspace &W)
{
    if (!outv || !inv || !pmin || !pmax || !sz || !P) return -1;
    if (sz[0] < 13 || sz[1] < 13 || sz[2] < 13 || bflag > 63) return -2;
    if (P->dissipation < 0 || P->dissipation > 2 || (P->dissipation && P->ccz4)) return -3;
    const Grid g{(int)sz[0], (int)sz[1], (int)sz[2]};
    const std::size_t n = (std::size_t)g.nx * g.ny * g.nz;
    const int nf = P->ccz4 ? 25 : 24;
    const double *in[NFMAX];
    double *out[NFMAX];
    for (int f = 0; f < nf; f++) {
        if (!inv[f] || !outv[f]) return -1;
        in[f] = inv[f] + offset;
        out[f] = outv[f] + offset;
    }
    const double h[3] = {(pmax[0] - pmin[0]) / (g.nx - 1), (pmax[1] - pmin[1]) / (g.ny - 1), (pmax[2] - pmin[2]) / (g.nz - 1)};
    W.resize(n, P->upwind != 0, nf);

    sweep_all(W, in, h, g, bflag, P->upwind != 0, nf);
    const double kappa1 = P->kappa[0], kappa2 = P->kappa[1], kappa3 = P->kappa[2];

    const double lambda0 = P->lambda[0], lambda1 = P->lambda[1], lambda2 = P->lambda[2], lambda3 = P->lambda[3];
    const double lambda_f0 = P->lambda_f[0], lambda_f1 = P->lambda_f[1];
    typedef double R;
    for (int k = 3; k < g.nz - 3; k++) {
        const double z = pmin[2] + k * h[2];
        for (int j = 3; j < g.ny - 3; j++) {
            const double y = pmin[1] + j * h[1];
            for (int i = 3; i < g.nx - 3; i++) {
                const double x = pmin[0] + i * h[0];
                const std::ptrdiff_t pp = g.idx(i, j, k);
                const double r_coord = std::sqrt(x * x + y * y + z * z);
                double eta = P->eta_const;
                if (P->eta_func) {
                    // GR/bssn.py:52 (USE_ETA_FUNC): R0 sqrt(igt^ij d_i chi d_j chi) / (1 - chi^p0)^p1
                    const double c0 = W.grad[1][0][pp], c1 = W.grad[1][1][pp], c2 = W.grad[1][2][pp];
                    eta = P->eta_func_r0 * std::sqrt(inverse_metric_quadratic(in, pp, c0, c1, c2)) *
                          std::pow(1.0 - std::pow(in[1][pp], P->eta_power[0]), -P->eta_power[1]);
                } else if (r_coord >= P->eta_r0) {
                    eta *= std::pow(P->eta_r0 / r_coord, P->eta_damping_exp);
                }
                const double xi1 = P->xi[1], xi2 = P->xi[2];
                const bool rochester = P->rochester != 0;
#define BSSN_U(f) in[f][pp]
#define BSSN_D1(f, d) W.grad[f][d][pp]
#define BSSN_D2(f, s) W.grad2[f][s][pp]
#define BSSN_ADV(f) ((in[6][pp] * W.agrad[f][0][pp] + in[7][pp] * W.agrad[f][1][pp]) + in[8][pp] * W.agrad[f][2][pp])
#define BSSN_LAP(f) metric_laplacian(in, W.grad2[f], pp)
// d_j (d_k beta^k) = (d_j d_0 beta^0 + d_j d_1 beta^1) + d_j d_2 beta^2
#define BSSN_GDB(j) ((W.grad2[6][SYM[j][0]][pp] + W.grad2[7][SYM[j][1]][pp]) + W.grad2[8][SYM[j][2]][pp])
// the Rochester gauge replaces the shift (6..8) and B (9..11) equations
#define BSSN_OUT(f, v) \
    if (!(rochester && (f) >= 6 && (f) <= 11)) out[f][pp] = (v)
#define BSSN_ALT_ON rochester
#define BSSN_OUT_ALT(f, v) out[f][pp] = (v)
#define BSSN_BARRIER()
                if (P->ccz4) {
#include "gen/ccz4_algebra.inc"
                } else {
#include "gen/bssn_algebra.inc"
                }
#undef BSSN_U
#undef BSSN_D1
#undef BSSN_D2
#undef BSSN_ADV
#undef BSSN_LAP
#undef BSSN_GDB
#undef BSSN_OUT
#undef BSSN_OUT_ALT
#undef BSSN_ALT_ON
#undef BSSN_BARRIER
            }
        }
    }

    // Sommerfeld outgoing-radiation condition on flagged faces (rhs.cpp:833-977),
    // applied after the pointwise stage and before dissipation.
    if (bflag) {
        const Box b = interior(g);
        for (int face = 0; face < 6; face++) {
            if (!(bflag & (1u << face))) continue;
            const int axis = face / 2;
            const int pos = (face & 1) ? g.n(axis) - 4 : 3;
            for (int f = 0; f < nf; f++) {
                for_box(g, b, axis, pos, [&](std::ptrdiff_t pp, int i, int j, int k) {
                    const double x = pmin[0] + i * h[0], y = pmin[1] + j * h[1], z = pmin[2] + k * h[2];
                    const double inv_r = 1.0 / std::sqrt(x * x + y * y + z * z);
                    out[f][pp] = -inv_r * (x * W.grad[f][0][pp] + y * W.grad[f][1][pp] + z * W.grad[f][2][pp] +
                                           BC_FALLOFF[f] * (in[f][pp] - BC_ASYM[f]));
                });
            }
        }
    }

    // Kreiss-Oliger (or TVB, rhs.cpp:729-757) sweeps into the first-derivative
    // arrays, then rhs += sigma (D_x + D_y + D_z) on the interior (rhs.cpp:324-369).
    const double sigma = P->ko_sigma;
    std::vector<double> lamv;
    double *lam[3] = {nullptr, nullptr, nullptr};
    if (P->dissipation) {
        lamv.assign(3 * n, 0.0);
        for (int d = 0; d < 3; d++) lam[d] = lamv.data() + d * n;
        spacetime_speeds(lam, in, n);
    }
    for (int f = 0; f < nf; f++) {
        for (int a = 0; a < 3; a++) {
            if (P->dissipation)
                tvb(W.grad[f][a], in[f], lam[a], h[a], g, bflag, a, P->dissipation == 1 ? 3 : 5);
            else
                kreiss_oliger(W.grad[f][a], in[f], h[a], g, bflag, a);
        }
        for_box(g, interior(g), 0, -1, [&](std::ptrdiff_t pp, int, int, int) {
            out[f][pp] += sigma * (W.grad[f][0][pp] + W.grad[f][1][pp] + W.grad[f][2][pp]);
        });
    }
    return 0;
}

}  // namespace

extern "C" int oracle_bssnrhs(double **out, const double **in, uint64_t offset, const double *pmin, const double *pmax,
                              const uint32_t *sz, uint32_t bflag, const oracle_params *p)
{
    Workspace W;
    return rhs_block(out, in, offset, pmin, pmax, sz, bflag, p, W);
}

extern "C" int oracle_bssnrhs_blocks(double **out, const double **in, uint32_t nblocks, const uint64_t *offsets,
                                     const double *pmin, const double *pmax, const uint32_t *sz, const uint32_t *bflag,
                                     const oracle_params *p, int nthreads)
{
    int err = 0;
    if (nthreads <= 0) nthreads = 1;
#pragma omp parallel num_threads(nthreads) reduction(| : err)
    {
        Workspace W;
#pragma omp for schedule(dynamic)
        for (int64_t b = 0; b < (int64_t)nblocks; b++)
            err |= rhs_block(out, in, offsets[b], pmin + 3 * b, pmax + 3 * b, sz + 3 * b, bflag[b], p, W) ? 1 : 0;
    }
    return err ? -1 : 0;
}

namespace {

// Hamiltonian and momentum constraints at the interior points of one block
// (GR/constraints.py:161-173), from the same derivative sweeps; with psi4 also
// the real and imaginary parts of the Weyl scalar Psi4 (GR/constraints.py:46-155)
// at the point's coordinates pmin + i h.
int constraints_impl(double **outv, const double **inv, uint64_t offset, const double *pmin, const double *pmax,
                     const uint32_t *sz, uint32_t bflag, bool psi4)
{
    if (!outv || !inv || !pmin || !pmax || !sz) return -1;
    if (sz[0] < 13 || sz[1] < 13 || sz[2] < 13 || bflag > 63) return -2;
    const Grid g{(int)sz[0], (int)sz[1], (int)sz[2]};
    const std::size_t n = (std::size_t)g.nx * g.ny * g.nz;
    const int nout = psi4 ? 6 : 4;
    const double *in[24];
    double *out[6];
    for (int f = 0; f < 24; f++) {
        if (!inv[f]) return -1;
        in[f] = inv[f] + offset;
    }
    for (int c = 0; c < nout; c++) {
        if (!outv[c]) return -1;
        out[c] = outv[c] + offset;
    }
    const double h[3] = {(pmax[0] - pmin[0]) / (g.nx - 1), (pmax[1] - pmin[1]) / (g.ny - 1), (pmax[2] - pmin[2]) / (g.nz - 1)};
    Workspace W;
    W.resize(n, false);
    sweep_all(W, in, h, g, bflag, false);
    typedef double R;
    for (int k = 3; k < g.nz - 3; k++)
        for (int j = 3; j < g.ny - 3; j++)
            for (int i = 3; i < g.nx - 3; i++) {
                const std::ptrdiff_t pp = g.idx(i, j, k);
#define BSSN_U(f) in[f][pp]
#define BSSN_D1(f, d) W.grad[f][d][pp]
#define BSSN_D2(f, s) W.grad2[f][s][pp]
#define BSSN_LAP(f) metric_laplacian(in, W.grad2[f], pp)
#define BSSN_OUT(c, v) out[c][pp] = (v)
#define BSSN_BARRIER()
                if (psi4) {
                    const R px = pmin[0] + i * h[0], py = pmin[1] + j * h[1], pz = pmin[2] + k * h[2];
#include "gen/bssn_constraints_psi4.inc"
                } else {
#include "gen/bssn_constraints.inc"
                }
#undef BSSN_U
#undef BSSN_D1
#undef BSSN_D2
#undef BSSN_LAP
#undef BSSN_OUT
#undef BSSN_BARRIER
            }
    return 0;
}

}  // namespace

extern "C" int oracle_constraints(double **outv, const double **inv, uint64_t offset, const double *pmin,
                                  const double *pmax, const uint32_t *sz, uint32_t bflag)
{
    return constraints_impl(outv, inv, offset, pmin, pmax, sz, bflag, false);
}

extern "C" int oracle_constraints_psi4(double **outv, const double **inv, uint64_t offset, const double *pmin,
                                       const double *pmax, const uint32_t *sz, uint32_t bflag)
{
    return constraints_impl(outv, inv, offset, pmin, pmax, sz, bflag, true);
}

extern "C" int oracle_deriv(int which, double *out, const double *u, double h, const uint32_t *sz, const double *beta,
                            uint32_t bflag, int upwind)
{
    if (!out || !u || !sz) return -1;
    const Grid g{(int)sz[0], (int)sz[1], (int)sz[2]};
    const int a = which % 3;
    switch (which / 3) {
    case 0: first(out, u, h, g, bflag, a); break;
    case 1: second(out, u, h, g, bflag, a); break;
    case 2:
        if (upwind) {
            if (!beta) return -1;
            advective(out, u, h, g, beta, bflag, a);
        } else {
            first(out, u, h, g, bflag, a);
        }
        break;
    case 3: kreiss_oliger(out, u, h, g, bflag, a); break;
    case 4:
    case 5:
        if (!beta) return -1;
        tvb(out, u, beta, h, g, bflag, a, which / 3 == 4 ? 3 : 5);
        break;
    default: return -1;
    }
    return 0;
}

extern "C" int oracle_spacetime_speeds(double **lam, const double **inv, uint64_t offset, const uint32_t *sz)
{
    if (!lam || !inv || !sz || !lam[0] || !lam[1] || !lam[2]) return -1;
    const double *in[24];
    for (int f = 0; f < 24; f++) {
        if (!inv[f]) return -1;
        in[f] = inv[f] + offset;
    }
    double *const l[3] = {lam[0] + offset, lam[1] + offset, lam[2] + offset};
    spacetime_speeds(l, in, (std::size_t)sz[0] * sz[1] * sz[2]);
    return 0;
}
