// 1-D finite-difference stencils with the reference's bflag face closures,
// evaluated from an LDS tile.  `u` points at the sample under evaluation,
// `s` is the stride along the sweep axis, `pos` the padded coordinate along
// that axis, `n` the padded size, `lo`/`hi` whether the block's low/high face
// on this axis is a domain boundary.
//
// Taps are applied left to right in the order bssn/src/derivs.cpp writes
// them, with contraction off, so each value is bit-equal to the reference
// sweep: 6th-order first derivative (deriv64_*, derivs.cpp:2055-2434),
// 6th-order second derivative (deriv64_xx.., 2443-2833), 4th-order
// shift-upwinded advective derivative (deriv42adv_*, 495-906) and
// Kreiss-Oliger dissipation (ko_deriv42_*, 915-1319).
#pragma once

#include "bssn_common.h"

namespace sgr {

#define SGR_EXACT _Pragma("clang fp contract(off)")

template <typename P>
__device__ __forceinline__ double tap(const P *u, int s, int o) { return (double)u[o * s]; }

// ---------------------------------------------------------------- first
template <typename P>
__device__ __forceinline__ double d1_int(const P *u, int s, const AxisConst &a)
{
    SGR_EXACT
    return (-1.0 * tap(u, s, -3) + 9.0 * tap(u, s, -2) - 45.0 * tap(u, s, -1) + 45.0 * tap(u, s, 1) -
            9.0 * tap(u, s, 2) + 1.0 * tap(u, s, 3)) * a.by60;
}

template <typename P>
__device__ __forceinline__ double d1_closure(const P *u, int s, int m, bool left, const AxisConst &a)
{
    SGR_EXACT
    if (left) {
        if (m == 0)
            return (-25.0 * tap(u, s, 0) + 48.0 * tap(u, s, 1) - 36.0 * tap(u, s, 2) + 16.0 * tap(u, s, 3) -
                    3.0 * tap(u, s, 4)) * a.by12;
        if (m == 1)
            return (-3.0 * tap(u, s, -1) - 10.0 * tap(u, s, 0) + 18.0 * tap(u, s, 1) - 6.0 * tap(u, s, 2) +
                    1.0 * tap(u, s, 3)) * a.by12;
        return (1.0 * tap(u, s, -2) - 8.0 * tap(u, s, -1) + 8.0 * tap(u, s, 1) - 1.0 * tap(u, s, 2)) * a.by12;
    }
    if (m == 0)
        return (1.0 * tap(u, s, -2) - 8.0 * tap(u, s, -1) + 8.0 * tap(u, s, 1) - 1.0 * tap(u, s, 2)) * a.by12;
    if (m == 1)
        return (-1.0 * tap(u, s, -3) + 6.0 * tap(u, s, -2) - 18.0 * tap(u, s, -1) + 10.0 * tap(u, s, 0) +
                3.0 * tap(u, s, 1)) * a.by12;
    return (3.0 * tap(u, s, -4) - 16.0 * tap(u, s, -3) + 36.0 * tap(u, s, -2) - 48.0 * tap(u, s, -1) +
            25.0 * tap(u, s, 0)) * a.by12;
}

template <typename P>
__device__ __forceinline__ double d1(const P *u, int s, int pos, int n, bool lo, bool hi, const AxisConst &a)
{
    if (lo && pos < GW + 3) return d1_closure(u, s, pos - GW, true, a);
    if (hi && pos >= n - GW - 3) return d1_closure(u, s, pos - (n - GW - 3), false, a);
    return d1_int(u, s, a);
}

// ---------------------------------------------------------------- second
template <typename P>
__device__ __forceinline__ double d2_int(const P *u, int s, const AxisConst &a)
{
    SGR_EXACT
    return (2.0 * tap(u, s, -3) - 27.0 * tap(u, s, -2) + 270.0 * tap(u, s, -1) - 490.0 * tap(u, s, 0) +
            270.0 * tap(u, s, 1) - 27.0 * tap(u, s, 2) + 2.0 * tap(u, s, 3)) * a.sq180;
}

template <typename P>
__device__ __forceinline__ double d2_closure(const P *u, int s, int m, bool left, const AxisConst &a)
{
    SGR_EXACT
    if (left) {
        if (m == 0)
            return (35.0 * tap(u, s, 0) - 104.0 * tap(u, s, 1) + 114.0 * tap(u, s, 2) - 56.0 * tap(u, s, 3) +
                    11.0 * tap(u, s, 4)) * a.sq12;
        if (m == 1)
            return (11.0 * tap(u, s, -1) - 20.0 * tap(u, s, 0) + 6.0 * tap(u, s, 1) + 4.0 * tap(u, s, 2) -
                    1.0 * tap(u, s, 3)) * a.sq12;
        return (-1.0 * tap(u, s, -2) + 16.0 * tap(u, s, -1) - 30.0 * tap(u, s, 0) + 16.0 * tap(u, s, 1) -
                1.0 * tap(u, s, 2)) * a.sq12;
    }
    if (m == 0)
        return (-1.0 * tap(u, s, -2) + 16.0 * tap(u, s, -1) - 30.0 * tap(u, s, 0) + 16.0 * tap(u, s, 1) -
                1.0 * tap(u, s, 2)) * a.sq12;
    if (m == 1)
        return (-1.0 * tap(u, s, -3) + 4.0 * tap(u, s, -2) + 6.0 * tap(u, s, -1) - 20.0 * tap(u, s, 0) +
                11.0 * tap(u, s, 1)) * a.sq12;
    return (11.0 * tap(u, s, -4) - 56.0 * tap(u, s, -3) + 114.0 * tap(u, s, -2) - 104.0 * tap(u, s, -1) +
            35.0 * tap(u, s, 0)) * a.sq12;
}

template <typename P>
__device__ __forceinline__ double d2(const P *u, int s, int pos, int n, bool lo, bool hi, const AxisConst &a)
{
    if (lo && pos < GW + 3) return d2_closure(u, s, pos - GW, true, a);
    if (hi && pos >= n - GW - 3) return d2_closure(u, s, pos - (n - GW - 3), false, a);
    return d2_int(u, s, a);
}

// ---------------------------------------------------------------- advective
template <typename P>
__device__ __forceinline__ double ad_int(const P *u, int s, bool pos_beta, const AxisConst &a)
{
    SGR_EXACT
    if (pos_beta)
        return (-3.0 * tap(u, s, -1) - 10.0 * tap(u, s, 0) + 18.0 * tap(u, s, 1) - 6.0 * tap(u, s, 2) +
                tap(u, s, 3)) * a.by12;
    return (-tap(u, s, -3) + 6.0 * tap(u, s, -2) - 18.0 * tap(u, s, -1) + 10.0 * tap(u, s, 0) +
            3.0 * tap(u, s, 1)) * a.by12;
}

template <typename P>
__device__ __forceinline__ double ad_fwd2(const P *u, int s, const AxisConst &a)
{
    SGR_EXACT
    return (-3.0 * tap(u, s, 0) + 4.0 * tap(u, s, 1) - tap(u, s, 2)) * a.by2;
}
template <typename P>
__device__ __forceinline__ double ad_cen2(const P *u, int s, const AxisConst &a)
{
    SGR_EXACT
    return (-tap(u, s, -1) + tap(u, s, 1)) * a.by2;
}
template <typename P>
__device__ __forceinline__ double ad_bwd2(const P *u, int s, const AxisConst &a)
{
    SGR_EXACT
    return (tap(u, s, -2) - 4.0 * tap(u, s, -1) + 3.0 * tap(u, s, 0)) * a.by2;
}

template <typename P>
__device__ __forceinline__ double adv(const P *u, int s, int pos, int n, bool lo, bool hi, double beta,
                                      const AxisConst &a)
{
    const bool bp = beta > 0.0;
    if (lo && pos < GW + 3) {
        const int m = pos - GW;
        if (m == 0) return ad_fwd2(u, s, a);
        if (m == 1) return bp ? ad_fwd2(u, s, a) : ad_cen2(u, s, a);
        return bp ? ad_int(u, s, true, a) : ad_bwd2(u, s, a);
    }
    if (hi && pos >= n - GW - 3) {
        const int m = pos - (n - GW - 3);
        if (m == 0) return bp ? ad_fwd2(u, s, a) : ad_int(u, s, false, a);
        if (m == 1) return bp ? ad_cen2(u, s, a) : ad_bwd2(u, s, a);
        return ad_bwd2(u, s, a);
    }
    return ad_int(u, s, bp, a);
}

// ---------------------------------------------------------------- Kreiss-Oliger
template <typename P>
__device__ __forceinline__ double ko_int(const P *u, int s, const AxisConst &a)
{
    SGR_EXACT
    return a.ko_pre * (-tap(u, s, -3) + 6.0 * tap(u, s, -2) - 15.0 * tap(u, s, -1) + 20.0 * tap(u, s, 0) -
                       15.0 * tap(u, s, 1) + 6.0 * tap(u, s, 2) - tap(u, s, 3));
}

template <typename P>
__device__ __forceinline__ double ko(const P *u, int s, int pos, int n, bool lo, bool hi, const AxisConst &a)
{
    SGR_EXACT
    if (lo && pos < GW + 3) {
        const int m = pos - GW;
        if (m == 0) return (tap(u, s, 3) - 3.0 * tap(u, s, 2) + 3.0 * tap(u, s, 1) - tap(u, s, 0)) / a.smr3;
        if (m == 1)
            return (tap(u, s, 3) - 6.0 * tap(u, s, 2) + 12.0 * tap(u, s, 1) - 10.0 * tap(u, s, 0) +
                    3.0 * tap(u, s, -1)) / a.smr2;
        return (tap(u, s, 3) - 6.0 * tap(u, s, 2) + 15.0 * tap(u, s, 1) - 19.0 * tap(u, s, 0) +
                12.0 * tap(u, s, -1) - 3.0 * tap(u, s, -2)) / a.smr1;
    }
    if (hi && pos >= n - GW - 3) {
        const int m = pos - (n - GW - 3);
        if (m == 0)
            return (tap(u, s, -3) - 6.0 * tap(u, s, -2) + 15.0 * tap(u, s, -1) - 19.0 * tap(u, s, 0) +
                    12.0 * tap(u, s, 1) - 3.0 * tap(u, s, 2)) / a.smr1;
        if (m == 1)
            return (tap(u, s, -3) - 6.0 * tap(u, s, -2) + 12.0 * tap(u, s, -1) - 10.0 * tap(u, s, 0) +
                    3.0 * tap(u, s, 1)) / a.smr2;
        return (tap(u, s, -3) - 3.0 * tap(u, s, -2) + 3.0 * tap(u, s, -1) - tap(u, s, 0)) / a.smr3;
    }
    return ko_int(u, s, a);
}

// ---------------------------------------------------------------- TVB
__device__ __forceinline__ double larger(double a, double b) { return a < b ? b : a; }  // std::max(a, b)

// TVB dissipation along one axis at one point (disstvb3_* / disstvb5_*,
// derivs.cpp:1348-2054).  l[o] (o = -4..3, l = L + 4) is the characteristic
// speed along the axis at tap o; only the taps the point's form reads need
// be set.  Points next to a flagged face take the Kreiss-Oliger closures
// (the reference writes the same formulas).  TVB3 takes its interior form at
// every other point (the reference leaves the first and last interior node
// before an unflagged face unwritten); TVB5 has one-sided forms at the first
// and last interior node.  pre: -1/(12h) for TVB3; 2/(75h) in x and y and
// -1/(64h) in z for TVB5, as the reference has them.
template <bool TVB5>
__device__ __forceinline__ double tvb(const double *u, int s, int pos, int n, bool lo, bool hi, const double *L,
                                      double pre, const AxisConst &a)
{
    SGR_EXACT
    if ((lo && pos < GW + 3) || (hi && pos >= n - GW - 3)) return ko(u, s, pos, n, lo, hi, a);
    const double *l = L + 4;
    if (!TVB5)
        return pre * (l[-2] * tap(u, s, -2) - 4.0 * l[-1] * tap(u, s, -1) + 6.0 * l[0] * tap(u, s, 0) -
                      4.0 * l[1] * tap(u, s, 1) + l[2] * tap(u, s, 2));
    if (pos == GW)
        return (l[3] * tap(u, s, 3) - 6.0 * l[2] * tap(u, s, 2) + 15.0 * l[1] * tap(u, s, 1) -
                19.0 * l[0] * tap(u, s, 0) + 12.0 * l[-1] * tap(u, s, -1) - 3.0 * l[-2] * tap(u, s, -2)) / a.smr1;
    if (pos == n - GW - 1)
        return (l[-4] * tap(u, s, -4) - 6.0 * l[-3] * tap(u, s, -3) + 15.0 * l[-2] * tap(u, s, -2) -
                19.0 * l[-1] * tap(u, s, -1) + 12.0 * l[0] * tap(u, s, 0) - 3.0 * l[1] * tap(u, s, 1)) / a.smr1;
    const double t3 = tap(u, s, 3), t2 = tap(u, s, 2), t1 = tap(u, s, 1), t0 = tap(u, s, 0);
    const double m1 = tap(u, s, -1), m2 = tap(u, s, -2), m3 = tap(u, s, -3);
    return pre * (larger(l[3], l[2]) * (t3 - t2) - 5.0 * larger(l[2], l[1]) * (t2 - t1) +
                  10.0 * larger(l[1], l[0]) * (t1 - t0) - 10.0 * larger(l[0], l[-1]) * (t0 - m1) +
                  5.0 * larger(l[-1], l[-2]) * (m1 - m2) - larger(l[-2], l[-3]) * (m2 - m3));
}

}  // namespace sgr
