// nuc_device.h -- device-side numerics for the NUC3d RHS sweeps on gfx950.
//
// Point formulas restate the reference schemes (file:line cited per function);
// the arithmetic is rearranged for the GPU (x*x instead of std::pow, one
// normalising division), which changes results only at the rounding level.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nucdev {

enum { WENO5JS = 0, WENO5Z = 1, WENO7JS = 2, CRWENO5 = 3, HCCS = 4 };
enum { VIS_NONE = 0, VIS_CD6 = 1, VIS_CD8 = 2 };
enum { RIE_LF = 0, RIE_AUSM = 1 };

template <typename T> __device__ __forceinline__ T sq(T x) { return x * x; }

// Interface stencil helpers.  A line buffer L is an LDS array valid on
// [-HALO, n+HALO); for interface m (between cells m-1 and m) the positive
// (f+) stencil in upwind coordinates is f[2+z] = L[m-1+z] and the negative
// one f[2+z] = L[m-z] (src/weno5js.cpp:262-275 vs :359-372).

// WENO5-JS, src/weno5js.cpp:62-111 (eps 1e-6, ideal weights 0.1/0.6/0.3).
template <typename T>
__device__ __forceinline__ T weno5js(T f0, T f1, T f2, T f3, T f4)
{
    const T g0 = T(13.0 / 12.0), g1 = T(0.25), eps = T(1.0e-6);
    T is0 = g0 * sq(f0 - T(2) * f1 + f2) + g1 * sq(f0 - T(4) * f1 + T(3) * f2);
    T is1 = g0 * sq(f1 - T(2) * f2 + f3) + g1 * sq(f1 - f3);
    T is2 = g0 * sq(f2 - T(2) * f3 + f4) + g1 * sq(T(3) * f2 - T(4) * f3 + f4);
    T q0 = T(1.0 / 3.0) * f0 + T(-7.0 / 6.0) * f1 + T(11.0 / 6.0) * f2;
    T q1 = T(-1.0 / 6.0) * f1 + T(5.0 / 6.0) * f2 + T(1.0 / 3.0) * f3;
    T q2 = T(1.0 / 3.0) * f2 + T(5.0 / 6.0) * f3 + T(-1.0 / 6.0) * f4;
    T a0 = T(0.1) / sq(eps + is0);
    T a1 = T(0.6) / sq(eps + is1);
    T a2 = T(0.3) / sq(eps + is2);
    return (a0 * q0 + a1 * q1 + a2 * q2) / (a0 + a1 + a2);
}

// WENO5-Z, src/weno5z.cpp:72-110 (eps 1e-40).  typo reproduces the positive
// inner sweep's smoothness indicator is2 (src/weno5z.cpp:80).
template <typename T>
__device__ __forceinline__ T weno5z(T f0, T f1, T f2, T f3, T f4, bool typo)
{
    const T g0 = T(13.0 / 12.0), g1 = T(0.25);
    const T eps = T(1.0e-40);
    T is0 = g0 * sq(f0 - T(2) * f1 + f2) + g1 * sq(f0 - T(4) * f1 + T(3) * f2);
    T is1 = g0 * sq(f1 - T(2) * f2 + f3) + g1 * sq(f1 - f3);
    T d2 = typo ? (f2 - f3) : (f2 - T(2) * f3 + f4);
    T is2 = g0 * sq(d2) + g1 * sq(T(3) * f2 - T(4) * f3 + f4);
    T q0 = T(1.0 / 3.0) * f0 + T(-7.0 / 6.0) * f1 + T(11.0 / 6.0) * f2;
    T q1 = T(-1.0 / 6.0) * f1 + T(5.0 / 6.0) * f2 + T(1.0 / 3.0) * f3;
    T q2 = T(1.0 / 3.0) * f2 + T(5.0 / 6.0) * f3 + T(-1.0 / 6.0) * f4;
    T tau = fabs(is2 - is0);
    T a0 = T(0.1) * (T(1) + sq(tau / (is0 + eps)));
    T a1 = T(0.6) * (T(1) + sq(tau / (is1 + eps)));
    T a2 = T(0.3) * (T(1) + sq(tau / (is2 + eps)));
    return (a0 * q0 + a1 * q1 + a2 * q2) / (a0 + a1 + a2);
}

// 7-point WENO with Z weights, src/weno7js.cpp:79-133 (eps 1e-20); hccs
// closures use the same formula with eps 1e-40 (src/hccs.cpp:85-135).
// The smoothness indicators are expanded quadratic forms whose coefficients
// (up to 17246) cancel to O(h^2) on smooth data, so any reassociation or FMA
// moves them by ~1e-6 relative and shifts the Z weights.  This function is
// therefore evaluated without contraction and in the reference's operation
// order, which makes it bit-identical to the reference for the indicators.
template <typename T>
__device__ __forceinline__ T weno7z(const T *f, T eps)
{
#pragma clang fp contract(off)
    T is0 = f[0] * (T(547) * f[0] - T(3882) * f[1] + T(4642) * f[2] - T(1854) * f[3])
          + f[1] * (T(7043) * f[1] - T(17246) * f[2] + T(7042) * f[3])
          + f[2] * (T(11003) * f[2] - T(9402) * f[3]) + T(2107) * f[3] * f[3];
    T is1 = f[1] * (T(267) * f[1] - T(1642) * f[2] + T(1602) * f[3] - T(494) * f[4])
          + f[2] * (T(2843) * f[2] - T(5966) * f[3] + T(1922) * f[4])
          + f[3] * (T(3443) * f[3] - T(2522) * f[4]) + T(547) * f[4] * f[4];
    T is2 = f[2] * (T(547) * f[2] - T(2522) * f[3] + T(1922) * f[4] - T(494) * f[5])
          + f[3] * (T(3443) * f[3] - T(5966) * f[4] + T(1602) * f[5])
          + f[4] * (T(2843) * f[4] - T(1642) * f[5]) + T(267) * f[5] * f[5];
    T is3 = f[3] * (T(2107) * f[3] - T(9402) * f[4] + T(7042) * f[5] - T(1854) * f[6])
          + f[4] * (T(11003) * f[4] - T(17246) * f[5] + T(4642) * f[6])
          + f[5] * (T(7043) * f[5] - T(3882) * f[6]) + T(547) * f[6] * f[6];
    T q0 = T(-3.0 / 12.0) * f[0] + T(13.0 / 12.0) * f[1] + T(-23.0 / 12.0) * f[2] + T(25.0 / 12.0) * f[3];
    T q1 = T(1.0 / 12.0) * f[1] + T(-5.0 / 12.0) * f[2] + T(13.0 / 12.0) * f[3] + T(3.0 / 12.0) * f[4];
    T q2 = T(-1.0 / 12.0) * f[2] + T(7.0 / 12.0) * f[3] + T(7.0 / 12.0) * f[4] + T(-1.0 / 12.0) * f[5];
    T q3 = T(3.0 / 12.0) * f[3] + T(13.0 / 12.0) * f[4] + T(-5.0 / 12.0) * f[5] + T(1.0 / 12.0) * f[6];
    T tau = fabs(is0 - is3);
    T a0 = T(1.0 / 35.0) * (T(1) + sq(tau / (eps + is0)));
    T a1 = T(12.0 / 35.0) * (T(1) + sq(tau / (eps + is1)));
    T a2 = T(18.0 / 35.0) * (T(1) + sq(tau / (eps + is2)));
    T a3 = T(4.0 / 35.0) * (T(1) + sq(tau / (eps + is3)));
    T asum = a0 + a1 + a2 + a3;
    return a0 / asum * q0 + a1 / asum * q1 + a2 / asum * q2 + a3 / asum * q3;
}

// CRWENO5 interior row in upwind coordinates, src/crweno5.cpp:133-162:
// lo*x(behind) + di*x(here) + hi*x(ahead) = rhs.
template <typename T>
__device__ __forceinline__ void crweno5_row(T f0, T f1, T f2, T f3, T f4, T &lo, T &di, T &hi, T &rhs)
{
    const T g0 = T(13.0 / 12.0), g1 = T(0.25), eps = T(1.0e-6);
    T is0 = g0 * sq(f0 - T(2) * f1 + f2) + g1 * sq(f0 - T(4) * f1 + T(3) * f2);
    T is1 = g0 * sq(f1 - T(2) * f2 + f3) + g1 * sq(f1 - f3);
    T is2 = g0 * sq(f2 - T(2) * f3 + f4) + g1 * sq(T(3) * f2 - T(4) * f3 + f4);
    T a0 = T(0.2) / sq(eps + is0);
    T a1 = T(0.5) / sq(eps + is1);
    T a2 = T(0.3) / sq(eps + is2);
    T r = T(1) / (a0 + a1 + a2);
    T w0 = a0 * r, w1 = a1 * r, w2 = a2 * r;
    lo = T(2.0 / 3.0) * w0 + T(1.0 / 3.0) * w1;
    di = T(1.0 / 3.0) * w0 + T(2.0 / 3.0) * (w1 + w2);
    hi = T(1.0 / 3.0) * w2;
    rhs = (w0 * f1 + (T(5) * (w0 + w1) + w2) * f2 + (w1 + T(5) * w2) * f3) * T(1.0 / 6.0);
}

// hccs interior row, src/hccs.cpp:138-183: Z-type CRWENO weights (eps 1e-40)
// blended with the 7th-order upwind compact scheme through theta.
template <typename T>
__device__ __forceinline__ void hccs_row(T f0, T f1, T f2, T f3, T f4, T &lo, T &di, T &hi, T &rhs)
{
    const T g0 = T(13.0 / 12.0), g1 = T(0.25);
    const T eps = T(1.0e-40);
    T is0 = g0 * sq(f0 - T(2) * f1 + f2) + g1 * sq(f0 - T(4) * f1 + T(3) * f2);
    T is1 = g0 * sq(f1 - T(2) * f2 + f3) + g1 * sq(f1 - f3);
    T is2 = g0 * sq(f2 - T(2) * f3 + f4) + g1 * sq(T(3) * f2 - T(4) * f3 + f4);
    T q7 = (-f0 + T(19) * f1 + T(239) * f2 + T(159) * f3 + T(4) * f4) / T(420);
    T tau = fabs(is2 - is0);
    T a0 = T(0.2) * (T(1) + sq(tau / (is0 + eps)));
    T a1 = T(0.5) * (T(1) + sq(tau / (is1 + eps)));
    T a2 = T(0.3) * (T(1) + sq(tau / (is2 + eps)));
    T asum = a0 + a1 + a2;
    T theta = T(1) / (T(1) + sq(asum - T(1)));
    if (theta > T(0.5)) {
        lo = T(2.0 / 7.0); di = T(4.0 / 7.0); hi = T(1.0 / 7.0); rhs = q7;
    } else {
        T r = T(1) / asum;
        T w0 = a0 * r, w1 = a1 * r, w2 = a2 * r;
        T s = (T(1) - theta) / theta;
        lo = (T(2.0 / 3.0) * w0 + T(1.0 / 3.0) * w1) * s + T(2.0 / 7.0);
        di = (T(1.0 / 3.0) * w0 + T(2.0 / 3.0) * (w1 + w2)) * s + T(4.0 / 7.0);
        hi = (T(1.0 / 3.0) * w2) * s + T(1.0 / 7.0);
        rhs = ((w0 * f1 + (T(5) * (w0 + w1) + w2) * f2 + (w1 + T(5) * w2) * f3) * T(1.0 / 6.0)) * s + q7;
    }
}

// ---------------------------------------------------------------------------
// Wave-level tridiagonal solve.  One wave (64 lanes) owns one line; lane k
// owns CPT consecutive rows.  Partition method: each lane eliminates inside
// its chunk keeping the dependence on its first/last unknowns (X0, X1); the
// reduced system is folded to one unknown per lane (X1) and solved by
// parallel cyclic reduction over the 64 lanes with shuffles; the interior
// follows by substitution.  Equivalent to the reference's Thomas sweep
// (src/crweno5.cpp:208-219) up to rounding.
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void pcr64(T &al, T &be, T &ga, T &de, int lane)
{
#pragma unroll
    for (int s = 1; s < 64; s <<= 1) {
        T alm = __shfl_up(al, s), bem = __shfl_up(be, s), gam = __shfl_up(ga, s), dem = __shfl_up(de, s);
        T alp = __shfl_down(al, s), bep = __shfl_down(be, s), gap = __shfl_down(ga, s), dep = __shfl_down(de, s);
        if (lane < s) { alm = T(0); bem = T(1); gam = T(0); dem = T(0); }
        if (lane + s >= 64) { alp = T(0); bep = T(1); gap = T(0); dep = T(0); }
        T k1 = al / bem;
        T k2 = ga / bep;
        T nal = -alm * k1;
        T nga = -gap * k2;
        T nbe = be - gam * k1 - alp * k2;
        T nde = de - dem * k1 - dep * k2;
        al = nal; be = nbe; ga = nga; de = nde;
    }
}

// a[r]*x[r-1] + b[r]*x[r] + c[r]*x[r+1] = d[r]; rows of lane k are
// k*CPT + r.  The first row of lane 0 and the last row of lane 63 must not
// couple outward (a = 0, c = 0).  Solution returned in d[].
template <int CPT, typename T>
__device__ __forceinline__ void tri_solve_wave(T *a, T *b, T *c, T *d, int lane)
{
    if constexpr (CPT == 1) {
        T al = a[0], be = b[0], ga = c[0], de = d[0];
        pcr64(al, be, ga, de, lane);
        d[0] = de / be;
    } else {
        // forward: rows 0,1 normalised; row r>=2 expressed with x0 (coef a')
        T ap[CPT], cp[CPT], dp[CPT];
        {
            T r0 = T(1) / b[0];
            ap[0] = a[0] * r0; cp[0] = c[0] * r0; dp[0] = d[0] * r0;
            T r1 = T(1) / b[1];
            ap[1] = a[1] * r1; cp[1] = c[1] * r1; dp[1] = d[1] * r1;
        }
#pragma unroll
        for (int r = 2; r < CPT; r++) {
            T rr = T(1) / (b[r] - a[r] * cp[r - 1]);
            dp[r] = rr * (d[r] - a[r] * dp[r - 1]);
            ap[r] = -rr * a[r] * ap[r - 1];
            cp[r] = rr * c[r];
        }
        // backward: rows CPT-2 .. 1 expressed with x0 and x_{CPT-1}
        T app[CPT], cpp[CPT], dpp[CPT];
        app[CPT - 2] = ap[CPT - 2]; cpp[CPT - 2] = cp[CPT - 2]; dpp[CPT - 2] = dp[CPT - 2];
#pragma unroll
        for (int r = CPT - 3; r >= 1; r--) {
            dpp[r] = dp[r] - cp[r] * dpp[r + 1];
            app[r] = ap[r] - cp[r] * app[r + 1];
            cpp[r] = -cp[r] * cpp[r + 1];
        }
        T A0, C0, D0;
        if constexpr (CPT == 2) {
            A0 = ap[0]; C0 = cp[0]; D0 = dp[0];
        } else {
            T r0 = T(1) / (T(1) - cp[0] * app[1]);
            A0 = r0 * ap[0];
            C0 = -r0 * cp[0] * cpp[1];
            D0 = r0 * (dp[0] - cp[0] * dpp[1]);
        }
        T A1 = ap[CPT - 1], C1 = cp[CPT - 1], D1 = dp[CPT - 1];
        // fold X0 out: tridiagonal system in X1 over lanes
        T A0n = __shfl_down(A0, 1), C0n = __shfl_down(C0, 1), D0n = __shfl_down(D0, 1);
        if (lane == 63) { A0n = T(0); C0n = T(0); D0n = T(0); }
        T al = -A1 * A0;
        T be = T(1) - A1 * C0 - C1 * A0n;
        T ga = -C1 * C0n;
        T de = D1 - A1 * D0 - C1 * D0n;
        pcr64(al, be, ga, de, lane);
        T X1 = de / be;
        T X1m = __shfl_up(X1, 1);
        if (lane == 0) X1m = T(0);
        T X0 = D0 - A0 * X1m - C0 * X1;
        d[0] = X0;
        d[CPT - 1] = X1;
#pragma unroll
        for (int r = 1; r <= CPT - 2; r++) d[r] = dpp[r] - app[r] * X0 - cpp[r] * X1;
    }
}

}  // namespace nucdev
