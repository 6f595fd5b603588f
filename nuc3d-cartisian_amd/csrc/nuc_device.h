// nuc_device.h -- device-side numerics for the NUC3d RHS sweeps on gfx950.
//
// Point formulas restate the reference schemes (file:line cited per function);
// the arithmetic is rearranged for the GPU (x*x instead of std::pow, one
// normalising division), which changes results only at the rounding level.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nucdev {

enum { WENO5JS = 0, WENO5Z = 1, WENO7JS = 2, CRWENO5 = 3, HCCS = 4, WENO3JS = 5, UPWIND1ST = 6, HCCS78 = 7, HCCS77 = 8 };
enum { VIS_NONE = 0, VIS_CD6 = 1, VIS_CD8 = 2, VIS_CD2 = 3 };
enum { RIE_LF = 0, RIE_AUSM = 1, RIE_AUSMP = 2 };   // AUSM+ runs the AUSM kernels with its pressure splitting

template <typename T> __device__ __forceinline__ T sq(T x) { return x * x; }

// WENO7 smoothness indicators as sums of squares, for FP32 stencils (the
// mixed path).  An exact LDL^T factorisation of each of the reference's
// quadratic forms (src/weno7js.cpp:80-99) in the differences g_i = f_i - f_{i+1}
// (the forms vanish on constants):
//   is0(f0..f3) = (547 g0 - 1394 g1 + 927 g2)^2 / 547
//               + 85040 / (547 * 2126^2) * (2126 g1 - 3167 g2)^2 + 203060 / 1063 * g2^2
//   is1(f0..f3) = (267 g0 - 554 g1 + 247 g2)^2 / 267
//               + 85040 / (267 * 2126^2) * (2126 g1 - 1345 g2)^2 + 203060 / 1063 * g2^2
// and is2, is3 are is1, is0 of the mirrored stencil.  Every term is
// non-negative, so FP32 evaluates them to FP32's relative precision instead
// of cancelling terms up to 17246 f^2.  outer: the is0 form, else is1's.
template <typename T>
__device__ __forceinline__ T w7_sos(T a, T b, T c, T d, bool outer)
{
    const T g0 = a - b, g1 = b - c, g2 = c - d;
    const T e0 = outer ? T(547) * g0 - T(1394) * g1 + T(927) * g2 : T(267) * g0 - T(554) * g1 + T(247) * g2;
    const T e1 = T(2126) * g1 - (outer ? T(3167) : T(1345)) * g2;
    const T k0 = outer ? T(1.0 / 547.0) : T(1.0 / 267.0);
    const T k1 = outer ? T(85040.0 / (547.0 * 2126.0 * 2126.0)) : T(85040.0 / (267.0 * 2126.0 * 2126.0));
    return k0 * e0 * e0 + k1 * e1 * e1 + T(203060.0 / 1063.0) * g2 * g2;
}

// 7-point WENO with Z weights, src/weno7js.cpp:79-133 (eps 1e-20); hccs
// closures use the same formula with eps 1e-40 (src/hccs.cpp:85-135), hccs78
// edges with 1e-20 (src/hccs78.cpp:246-304).
// FP64 stencils: the smoothness indicators are the reference's expanded
// quadratic forms, whose coefficients (up to 17246) cancel to O(h^2) on smooth
// data, so any reassociation or FMA moves them by ~1e-6 relative and shifts
// the Z weights.  They are therefore evaluated without contraction and in the
// reference's operation order, which makes them bit-identical to the
// reference.  FP32 stencils (the mixed path) use the sums of squares above.
// T: indicators and candidate values; TW: the nonlinear weights (the Z
// weights square tau/(eps+is) with eps 1e-20/1e-40, outside FP32's range).
// asum receives the sum of the unnormalised weights (hccs78's blend factor).
template <typename T, typename TW>
__device__ __forceinline__ TW weno7z(const T *f, TW eps, TW &asum)
{
#pragma clang fp contract(off)
    T is0, is1, is2, is3;
    if constexpr (sizeof(T) == 4) {
        is0 = w7_sos(f[0], f[1], f[2], f[3], true);
        is1 = w7_sos(f[1], f[2], f[3], f[4], false);
        is2 = w7_sos(f[5], f[4], f[3], f[2], false);
        is3 = w7_sos(f[6], f[5], f[4], f[3], true);
    } else {
        is0 = f[0] * (T(547) * f[0] - T(3882) * f[1] + T(4642) * f[2] - T(1854) * f[3])
            + f[1] * (T(7043) * f[1] - T(17246) * f[2] + T(7042) * f[3])
            + f[2] * (T(11003) * f[2] - T(9402) * f[3]) + T(2107) * f[3] * f[3];
        is1 = f[1] * (T(267) * f[1] - T(1642) * f[2] + T(1602) * f[3] - T(494) * f[4])
            + f[2] * (T(2843) * f[2] - T(5966) * f[3] + T(1922) * f[4])
            + f[3] * (T(3443) * f[3] - T(2522) * f[4]) + T(547) * f[4] * f[4];
        is2 = f[2] * (T(547) * f[2] - T(2522) * f[3] + T(1922) * f[4] - T(494) * f[5])
            + f[3] * (T(3443) * f[3] - T(5966) * f[4] + T(1602) * f[5])
            + f[4] * (T(2843) * f[4] - T(1642) * f[5]) + T(267) * f[5] * f[5];
        is3 = f[3] * (T(2107) * f[3] - T(9402) * f[4] + T(7042) * f[5] - T(1854) * f[6])
            + f[4] * (T(11003) * f[4] - T(17246) * f[5] + T(4642) * f[6])
            + f[5] * (T(7043) * f[5] - T(3882) * f[6]) + T(547) * f[6] * f[6];
    }
    T q0 = T(-3.0 / 12.0) * f[0] + T(13.0 / 12.0) * f[1] + T(-23.0 / 12.0) * f[2] + T(25.0 / 12.0) * f[3];
    T q1 = T(1.0 / 12.0) * f[1] + T(-5.0 / 12.0) * f[2] + T(13.0 / 12.0) * f[3] + T(3.0 / 12.0) * f[4];
    T q2 = T(-1.0 / 12.0) * f[2] + T(7.0 / 12.0) * f[3] + T(7.0 / 12.0) * f[4] + T(-1.0 / 12.0) * f[5];
    T q3 = T(3.0 / 12.0) * f[3] + T(13.0 / 12.0) * f[4] + T(-5.0 / 12.0) * f[5] + T(1.0 / 12.0) * f[6];
    const TW b0 = is0, b1 = is1, b2 = is2, b3 = is3;
    TW tau = fabs(b0 - b3);
    TW a0 = TW(1.0 / 35.0) * (TW(1) + sq(tau / (eps + b0)));
    TW a1 = TW(12.0 / 35.0) * (TW(1) + sq(tau / (eps + b1)));
    TW a2 = TW(18.0 / 35.0) * (TW(1) + sq(tau / (eps + b2)));
    TW a3 = TW(4.0 / 35.0) * (TW(1) + sq(tau / (eps + b3)));
    asum = a0 + a1 + a2 + a3;
    return a0 / asum * TW(q0) + a1 / asum * TW(q1) + a2 / asum * TW(q2) + a3 / asum * TW(q3);
}

template <typename T, typename TW>
__device__ __forceinline__ TW weno7z(const T *f, TW eps)
{
    TW asum;
    return weno7z<T, TW>(f, eps, asum);
}

// hccs78 inner point, src/hccs78.cpp:66-117: the WENO7-Z value (eps 1e-20)
// blended with the 8-point central value through theta = 1/(1 + (asum-1)^2).
// f[0..7]: one downwind point more than WENO7.
template <typename T, typename TW>
__device__ __forceinline__ TW hccs78_inner(const T *f)
{
    TW asum;
    TW w = weno7z<T, TW>(f, TW(1.0e-20), asum);
    TW q8 = (TW(-3) * (TW(f[7]) + TW(f[0])) + TW(29) * (TW(f[6]) + TW(f[1])) - TW(139) * (TW(f[5]) + TW(f[2]))
             + TW(533) * (TW(f[4]) + TW(f[3]))) / TW(840);
    TW theta = TW(1) / (TW(1) + sq(asum - TW(1)));
    return (TW(1) - theta) * w + theta * q8;
}

// hccs77 interior row, src/hccs77.cpp:86-163: WENO7-Z (eps 1e-20) and the
// 7th-order upwind compact row blended through theta = 1/(1 + (asum-1)^2);
// theta > 0.5 keeps the compact row alone.  f[0..6]; TW: weights and the row.
template <typename T, typename TW>
__device__ __forceinline__ void hccs77_row(const T *f, TW &lo, TW &di, TW &hi, TW &rhs)
{
    TW asum;
    TW w = weno7z<T, TW>(f, TW(1.0e-20), asum);
    TW q7 = (-TW(f[1]) + TW(19) * TW(f[2]) + TW(239) * TW(f[3]) + TW(159) * TW(f[4]) + TW(4) * TW(f[5])) / TW(420);
    TW theta = TW(1) / (TW(1) + sq(asum - TW(1)));
    if (theta > TW(0.5)) {
        lo = TW(2.0 / 7.0); di = TW(4.0 / 7.0); hi = TW(1.0 / 7.0); rhs = q7;
    } else {
        lo = TW(2.0 / 7.0) * theta;
        di = TW(1) + (TW(4.0 / 7.0) - TW(1)) * theta;
        hi = TW(1.0 / 7.0) * theta;
        rhs = w * (TW(1) - theta) + q7 * theta;
    }
}

// WENO3-JS, src/weno3js.cpp:80-93 (eps 1e-6): f[0..2] upwind
template <typename T, typename TW>
__device__ __forceinline__ TW weno3js(const T *f)
{
    T is0 = sq(f[0] - f[1]), is1 = sq(f[1] - f[2]);
    T q0 = T(-0.5) * f[0] + T(1.5) * f[1];
    T q1 = T(0.5) * f[1] + T(0.5) * f[2];
    TW a0 = TW(1.0 / 3.0) / sq(TW(1.0e-6) + TW(is0));
    TW a1 = TW(2.0 / 3.0) / sq(TW(1.0e-6) + TW(is1));
    return (a0 * TW(q0) + a1 * TW(q1)) / (a0 + a1);
}

// hccs interior row, src/hccs.cpp:138-183: Z-type CRWENO weights (eps 1e-40)
// blended with the 7th-order upwind compact scheme through theta.
// T: indicators and stencil values; TW: weights and the row.
template <typename T, typename TW>
__device__ __forceinline__ void hccs_row(T f0, T f1, T f2, T f3, T f4, TW &lo, TW &di, TW &hi, TW &rhs)
{
    const T g0 = T(13.0 / 12.0), g1 = T(0.25);
    const TW eps = TW(1.0e-40);
    T is0 = g0 * sq(f0 - T(2) * f1 + f2) + g1 * sq(f0 - T(4) * f1 + T(3) * f2);
    T is1 = g0 * sq(f1 - T(2) * f2 + f3) + g1 * sq(f1 - f3);
    T is2 = g0 * sq(f2 - T(2) * f3 + f4) + g1 * sq(T(3) * f2 - T(4) * f3 + f4);
    T q7 = (-f0 + T(19) * f1 + T(239) * f2 + T(159) * f3 + T(4) * f4) / T(420);
    const TW b0 = is0, b1 = is1, b2 = is2;
    TW tau = fabs(b2 - b0);
    TW a0 = TW(0.2) * (TW(1) + sq(tau / (b0 + eps)));
    TW a1 = TW(0.5) * (TW(1) + sq(tau / (b1 + eps)));
    TW a2 = TW(0.3) * (TW(1) + sq(tau / (b2 + eps)));
    TW asum = a0 + a1 + a2;
    TW theta = TW(1) / (TW(1) + sq(asum - TW(1)));
    if (theta > TW(0.5)) {
        lo = TW(2.0 / 7.0); di = TW(4.0 / 7.0); hi = TW(1.0 / 7.0); rhs = q7;
    } else {
        TW r = TW(1) / asum;
        TW w0 = a0 * r, w1 = a1 * r, w2 = a2 * r;
        TW s = (TW(1) - theta) / theta;
        TW F1 = f1, F2 = f2, F3 = f3;
        lo = (TW(2.0 / 3.0) * w0 + TW(1.0 / 3.0) * w1) * s + TW(2.0 / 7.0);
        di = (TW(1.0 / 3.0) * w0 + TW(2.0 / 3.0) * (w1 + w2)) * s + TW(4.0 / 7.0);
        hi = (TW(1.0 / 3.0) * w2) * s + TW(1.0 / 7.0);
        rhs = ((w0 * F1 + (TW(5) * (w0 + w1) + w2) * F2 + (w1 + TW(5) * w2) * F3) * TW(1.0 / 6.0)) * s + TW(q7);
    }
}

}  // namespace nucdev
