// nuc_lines.h -- per-line interface computation shared by the production
// sweep kernels and the per-Field operator kernels.
//
// One wave owns one line held in LDS on [-HALO, 32R+HALO).  Lanes 0..31
// reconstruct f^+ and lanes 32..63 reconstruct f^- (src/euler3d.cpp:334-387
// does the two as separate sweeps).  Lane h of a half owns a chunk of R
// consecutive interfaces; the f^- half takes the chunks in mirrored order
// (lane h <-> chunk 31-h, rows reversed), which turns the f^- stencils and its
// tridiagonal system into the same shape as f^+, so both halves execute one
// instruction stream.
//
// Rank-local zones follow the reference's inner/boundary split per MPI block
// of length nloc (src/fieldOperator.cpp:387-465):
//   explicit boundary interfaces  ml in [0, ts-1] U [nloc-ts+1, nloc]
//   inner interfaces              ml in [ts, nloc-ts]
//   compact schemes: ml = ts and ml = nloc-ts are explicit closure rows,
//   ts < ml < nloc-ts are tridiagonal rows (src/crweno5.cpp:83-206).
#pragma once
#include "nuc_device.h"

namespace nucdev {

constexpr int HALO = 4;

struct Zone {
    int nrows;    // interfaces on this line (n periodic, n+1 with ghost strips)
    int nloc;     // rank-local line length
    int ts;       // inviscid tilesize of the scheme
    int off = 0;  // position of interface 0 on the full line (a segment of a long line)
};

// reciprocal: hardware estimate refined by two Newton steps (|err| <= 1 ulp)
__device__ __forceinline__ double frcp(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}
// FP32 (mixed path): the hardware reciprocal, 1 ulp
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }

// Unnormalised WENO5 weights n_k ~ c_k / (eps + is_k)^2.  FP64 uses the
// division-free products c_k * s_i * s_j (one reciprocal per row in total);
// FP32 uses c_k * rcp(s_k), since the products can leave FP32's range.
template <typename TW>
__device__ __forceinline__ void w5_weights(TW b0, TW b1, TW b2, TW c0, TW c1, TW c2, TW &n0, TW &n1, TW &n2)
{
    const TW eps = TW(1.0e-6);
    TW s0 = sq(eps + b0), s1 = sq(eps + b1), s2 = sq(eps + b2);
    if constexpr (sizeof(TW) == 8) {
        n0 = c0 * (s1 * s2); n1 = c1 * (s0 * s2); n2 = c2 * (s0 * s1);
    } else {
        n0 = c0 * frcp(s0); n1 = c1 * frcp(s1); n2 = c2 * frcp(s2);
    }
}

// upwind reach of a scheme's stencil, and the points it reads beyond the
// symmetric window (hccs78's 8-point central part)
template <int SCH> struct Radius {
    static constexpr int value = (SCH == WENO7JS || SCH == HCCS || SCH == HCCS78 || SCH == HCCS77) ? 3
                               : (SCH == WENO3JS || SCH == UPWIND1ST) ? 1 : 2;
};
template <int SCH> struct Extra { static constexpr int value = (SCH == HCCS78) ? 1 : 0; };

// Skewed LDS line layout: position y (y >= -B) lives at slot (y+B) + (y+B)/R,
// one pad slot after every R positions, B = HALO+1 rounded up to a multiple of
// R.  The f^+ line stores cell x at y = x; the f^- line is stored mirrored,
// y = 32R-2-x, so that both half waves read their windows upwards from
// y = hR-RAD-1: the lane's window then starts at slot h*(R+1) + at(-RAD-1)
// and every window element is a compile-time offset from one base.  Lanes R
// positions apart sit R+1 slots apart (odd), so chunk reads and the
// consecutive-cell writes of the load phase hit distinct LDS banks.
// pitch(): line stride in slots, = 4 (mod 16): the y/z load mapping writes 4
// lines x 4 consecutive cells per 16 lanes, which then hit distinct banks
// (measured: y/z sweep bank-conflict cycles -79% against a pitch = 2 mod 16).
template <int R> struct Skew {
    static constexpr int B = ((HALO + 1 + R - 1) / R) * R;
    __host__ __device__ static constexpr int at(int y) { return (y + B) + (y + B) / R; }
    __device__ __forceinline__ static int mir(int x) { return at(32 * R - 2 - x); }
    // slots up to cell ncap+HALO: the last one is a write-only dummy (halo
    // writes of interior cells go there instead of branching)
    static constexpr int span(int ncap) { return at(ncap + HALO) + 1; }
    static constexpr int pitch(int ncap) { return span(ncap) + ((4 - span(ncap) % 16) + 16) % 16; }
};

// pitch of an unskewed line of ncap cells with HALO ghosts on each side and
// one dummy slot, = 4 (mod 16)
__host__ __device__ constexpr int plain_pitch(int ncap)
{
    return (ncap + 2 * HALO + 1) + ((4 - (ncap + 2 * HALO + 1) % 16) + 16) % 16;
}

// Distance-1 lane shifts as DPP moves (no LDS round trip): shr1 gives lane i
// the value of lane i-1, shl1 the value of lane i+1; the lanes without a source
// (0, resp. 63) read 0.  Across the boundary of the two half waves the values
// do move (lane 32 reads lane 31), so callers mask h == 0 / h == 31 as they
// do for the width-32 shuffles.
__device__ __forceinline__ int dpp_shr1(int x) { return __builtin_amdgcn_update_dpp(0, x, 0x138, 0xF, 0xF, true); }
__device__ __forceinline__ int dpp_shl1(int x) { return __builtin_amdgcn_update_dpp(0, x, 0x130, 0xF, 0xF, true); }
__device__ __forceinline__ double shr1(double x)
{
    return __hiloint2double(dpp_shr1(__double2hiint(x)), dpp_shr1(__double2loint(x)));
}
__device__ __forceinline__ double shl1(double x)
{
    return __hiloint2double(dpp_shl1(__double2hiint(x)), dpp_shl1(__double2loint(x)));
}
__device__ __forceinline__ float shr1(float x) { return __int_as_float(dpp_shr1(__float_as_int(x))); }
__device__ __forceinline__ float shl1(float x) { return __int_as_float(dpp_shl1(__float_as_int(x))); }

// Distance-16 exchange inside each half wave with one v_permlane16_swap per
// dword: swapping a value with itself leaves, in the first result, the lanes of
// the odd 16-lane rows holding the value of the lane 16 below (lane i-16), and,
// in the second, the lanes of the even rows holding the value of the lane 16
// above (lane i+16).  up16 is meaningful for h >= 16, dn16 for h < 16 (h = lane
// index within the half wave); the other lanes are masked by the callers.
__device__ __forceinline__ void swap16(unsigned x, unsigned &up, unsigned &dn)
{
    auto r = __builtin_amdgcn_permlane16_swap(x, x, false, false);
    up = r[0]; dn = r[1];
}
__device__ __forceinline__ void swap16(double x, double &up, double &dn)
{
    unsigned hu, hd, lu, ld;
    swap16((unsigned)__double2hiint(x), hu, hd);
    swap16((unsigned)__double2loint(x), lu, ld);
    up = __hiloint2double((int)hu, (int)lu);
    dn = __hiloint2double((int)hd, (int)ld);
}
__device__ __forceinline__ void swap16(float x, float &up, float &dn)
{
    unsigned u, d;
    swap16((unsigned)__float_as_int(x), u, d);
    up = __int_as_float((int)u); dn = __int_as_float((int)d);
}

// An empty asm that takes x and gives it back: the value must exist at this
// point, and code that uses the result cannot move above it (volatile asm
// statements keep their order).  Used to evaluate a lane's rows in order.
template <typename T> __device__ __forceinline__ void pin(T &x) { asm volatile("" : "+v"(x)); }

// normalised parallel cyclic reduction over the 32 lanes of a half wave:
// row k: x_k + al*x_{k-1} + ga*x_{k+1} = de  ->  x_k = de on return
template <typename T>
__device__ __forceinline__ void pcr32(T &al, T &ga, T &de, int h)
{
#pragma unroll
    for (int s = 1; s < 32; s <<= 1) {
        T alm, gam, dem, alp, gap, dep;
        if (s == 1) {
            alm = shr1(al); gam = shr1(ga); dem = shr1(de);
            alp = shl1(al); gap = shl1(ga); dep = shl1(de);
        } else if (s == 16) {
            swap16(al, alm, alp); swap16(ga, gam, gap); swap16(de, dem, dep);
        } else {
            alm = __shfl_up(al, s, 32); gam = __shfl_up(ga, s, 32); dem = __shfl_up(de, s, 32);
            alp = __shfl_down(al, s, 32); gap = __shfl_down(ga, s, 32); dep = __shfl_down(de, s, 32);
        }
        // no masks: lanes h < s hold al = 0 and lanes h + s >= 32 hold ga = 0
        // exactly (the first and last rows do not couple outward, and each
        // level doubles those ranges), so the values they received from the
        // wrapped or neighbouring lanes, all finite, are multiplied by zero
        T r = frcp(T(1) - al * gam - ga * alp);
        T nal = -r * al * alm;
        T nga = -r * ga * gap;
        T nde = r * (de - al * dem - ga * dep);
        al = nal; ga = nga; de = nde;
    }
}

// Forward-eliminated rows of one lane's chunk (rows normalised):
//   r = 0: Af x_{-1} + x_0 + Cf x_1 = Df ;  r >= 1: Af x_0 + x_r + Cf x_{r+1} = Df
template <int R, typename T> struct Chunk {
    T Af[R], Cf[R], Df[R];
    __device__ __forceinline__ void row(int r, T a, T b, T c, T d)
    {
        if (r <= 1) {
            T rb = frcp(b);
            Af[r] = a * rb; Cf[r] = c * rb; Df[r] = d * rb;
        } else {
            T rr = frcp(b - a * Cf[r - 1]);
            Df[r] = rr * (d - a * Df[r - 1]);
            Af[r] = -rr * a * Af[r - 1];
            Cf[r] = rr * c;
        }
    }
    // Partition solve across the half wave (lane h holds chunk h in system
    // order; the first row of lane 0 and the last row of lane 31 must not
    // couple outward).  Solution returned in Df.  Equivalent to the
    // reference's Thomas sweep (src/crweno5.cpp:208-219) up to rounding.
    __device__ __forceinline__ void solve(int h)
    {
#pragma unroll
        for (int r = R - 3; r >= 1; r--) {
            T tD = Df[r] - Cf[r] * Df[r + 1];
            T tA = Af[r] - Cf[r] * Af[r + 1];
            Cf[r] = -Cf[r] * Cf[r + 1];
            Df[r] = tD; Af[r] = tA;
        }
        T A0, C0, D0;
        if constexpr (R == 2) {
            A0 = Af[0]; C0 = Cf[0]; D0 = Df[0];
        } else {
            T r0 = frcp(T(1) - Cf[0] * Af[1]);
            A0 = r0 * Af[0];
            C0 = -r0 * Cf[0] * Cf[1];
            D0 = r0 * (Df[0] - Cf[0] * Df[1]);
        }
        T A1 = Af[R - 1], C1 = Cf[R - 1], D1 = Df[R - 1];
        T A0n = shl1(A0), C0n = shl1(C0), D0n = shl1(D0);
        if (h == 31) { A0n = T(0); C0n = T(0); D0n = T(0); }
        T rb = frcp(T(1) - A1 * C0 - C1 * A0n);
        T al = -rb * A1 * A0;
        T ga = -rb * C1 * C0n;
        T de = rb * (D1 - A1 * D0 - C1 * D0n);
        // the reduced system's first and last rows do not couple outward
        if (h == 0) al = T(0);
        if (h == 31) ga = T(0);
        pcr32(al, ga, de, h);
        T X1 = de;
        T X1m = shr1(X1);
        if (h == 0) X1m = T(0);
        T X0 = D0 - A0 * X1m - C0 * X1;
#pragma unroll
        for (int r = 1; r <= R - 2; r++) Df[r] = Df[r] - Af[r] * X0 - Cf[r] * X1;
        Df[0] = X0;
        Df[R - 1] = X1;
    }
};

// WENO5 smoothness indicators and candidate values of a 5-point upwind stencil
// f0..f4 (src/weno5js.cpp, src/crweno5.cpp:87-129), computed from the
// stencil's differences d_j = f_j - f_{j-1} (j = 1..4) and second-difference
// terms e_k = 13/12 (d_{k+1} - d_k)^2 (k = 1..3), which a lane's consecutive
// rows share: f0-4f1+3f2 = 3d2-d1, f1-f3 = -(d2+d3), 3f2-4f3+f4 = d4-3d3,
// q0 = f2+(5d2-2d1)/6, q1 = f2+(d2+2d3)/6, q2 = f3-(2d3+d4)/6.  The linear
// parts are halved (h_k), so that b_k = e_k + h_k^2 with the 1/4 folded in.
template <typename T>
__device__ __forceinline__ void w5_diff(T f2, T f3, T d1, T d2, T d3, T d4, T e1, T e2, T e3, T &b0, T &b1, T &b2,
                                        T &q0, T &q1, T &q2, T &h2)
{
    const T h0 = T(1.5) * d2 - T(0.5) * d1, h1 = T(0.5) * (d2 + d3);
    h2 = T(0.5) * d4 - T(1.5) * d3;
    b0 = e1 + h0 * h0;
    b1 = e2 + h1 * h1;
    b2 = e3 + h2 * h2;
    q0 = f2 + (T(5.0 / 6.0) * d2 - T(1.0 / 3.0) * d1);
    q1 = f2 + (T(1.0 / 6.0) * d2 + T(1.0 / 3.0) * d3);
    q2 = f3 - (T(1.0 / 3.0) * d3 + T(1.0 / 6.0) * d4);
}

// Reconstruct one sign of a line.  Lw points at the lane's window in the
// skewed line buffer (slot h*(R+1); the f^- buffer is mirrored, see Skew):
// win[t] = Lw[at(t-RAD-1)] in upwind order, row r uses win[r .. r+2*RAD] in
// upwind coordinates.  m0r is the interface index of row 0 and mstep its step
// (+1 for f^+, -1 for f^- whose rows run backwards).
// T is the type of the stencils (smoothness indicators, candidate values);
// TW the type of the nonlinear weights and compact rows; TS the type of the
// tridiagonal solve and of the result (FP64 in both precision paths).
// full: every row of every lane is an interface of the line (n = 32R).
template <int SCH, int R, typename T, typename TW, typename TS>
__device__ __forceinline__ void line_half(const T *Lw, int m0r, int mstep, const Zone &z, int h, bool neg,
                                          TS out[R], bool full = false)
{
    constexpr int RAD = Radius<SCH>::value;
    constexpr int W = R + 2 * RAD + Extra<SCH>::value;
    T win[W];
#pragma unroll
    for (int t = 0; t < W; t++) win[t] = Lw[Skew<R>::at(t - RAD - 1)];
    // rank-local position of each row, without integer division in the loop
    int ml[R];
    {
        int m = m0r;
        int l = (m + z.off) % z.nloc;
#pragma unroll
        for (int r = 0; r < R; r++) {
            ml[r] = l;
            l += mstep;
            if (l >= z.nloc) l -= z.nloc;
            if (l < 0) l += z.nloc;
        }
    }
    if constexpr (SCH == WENO5JS) {
        // rows in order through a rolling window of differences (see CRWENO5)
        T f2 = win[2], f3 = win[3];
        T d1 = win[1] - win[0], d2 = f2 - win[1], d3 = f3 - f2;
        T e1 = T(13.0 / 12.0) * sq(d2 - d1), e2 = T(13.0 / 12.0) * sq(d3 - d2);
#pragma unroll
        for (int r = 0; r < R; r++) {
            T f4 = win[r + 4];
            T d4 = f4 - f3;
            T e3 = T(13.0 / 12.0) * sq(d4 - d3);
            T b0, b1, b2, q0, q1, q2, h2;
            w5_diff<T>(f2, f3, d1, d2, d3, d4, e1, e2, e3, b0, b1, b2, q0, q1, q2, h2);
            // alpha_k = c_k / s_k, normalised with a single reciprocal
            TW n0, n1, n2;
            w5_weights<TW>(TW(b0), TW(b1), TW(b2), TW(0.1), TW(0.6), TW(0.3), n0, n1, n2);
            TS o = TS((n0 * TW(q0) + n1 * TW(q1) + n2 * TW(q2)) * frcp(n0 + n1 + n2));
            pin(o); pin(f3); pin(f4); pin(d3); pin(d4); pin(e2); pin(e3);
            out[r] = o;
            f2 = f3; f3 = f4; d1 = d2; d2 = d3; d3 = d4; e1 = e2; e2 = e3;
        }
    } else if constexpr (SCH == WENO5Z) {
        // rows in order through a rolling window of differences (see CRWENO5)
        T f2 = win[2], f3 = win[3];
        T d1 = win[1] - win[0], d2 = f2 - win[1], d3 = f3 - f2;
        T e1 = T(13.0 / 12.0) * sq(d2 - d1), e2 = T(13.0 / 12.0) * sq(d3 - d2);
#pragma unroll
        for (int r = 0; r < R; r++) {
            T f4 = win[r + 4];
            T d4 = f4 - f3;
            T e3 = T(13.0 / 12.0) * sq(d4 - d3);
            T b0, b1, b2, q0, q1, q2, h2;
            w5_diff<T>(f2, f3, d1, d2, d3, d4, e1, e2, e3, b0, b1, b2, q0, q1, q2, h2);
            // the positive inner sweep's is2 reads f_{i+1} twice (src/weno5z.cpp:80):
            // 13/12 (f2 - f3)^2 in place of its second-difference term
            bool typo = !neg && ml[r] >= z.ts && ml[r] <= z.nloc - z.ts;
            if (typo) b2 = T(13.0 / 12.0) * sq(d3) + h2 * h2;
            const TW eps = TW(1.0e-40);
            const TW B0 = b0, B1 = b1, B2 = b2;
            TW tau = fabs(B2 - B0);
            TW a0 = TW(0.1) * (TW(1) + sq(tau / (B0 + eps)));
            TW a1 = TW(0.6) * (TW(1) + sq(tau / (B1 + eps)));
            TW a2 = TW(0.3) * (TW(1) + sq(tau / (B2 + eps)));
            TS o = TS((a0 * TW(q0) + a1 * TW(q1) + a2 * TW(q2)) * frcp(a0 + a1 + a2));
            pin(o); pin(f3); pin(f4); pin(d3); pin(d4); pin(e2); pin(e3);
            out[r] = o;
            f2 = f3; f3 = f4; d1 = d2; d2 = d3; d3 = d4; e1 = e2; e2 = e3;
        }
    } else if constexpr (SCH == WENO7JS) {
        // rows in order through a rolling 7-point window (see CRWENO5 below)
        T fw[7];
#pragma unroll
        for (int j = 0; j < 6; j++) fw[j] = win[j];
#pragma unroll
        for (int r = 0; r < R; r++) {
            fw[6] = win[r + 6];
            TS o = TS(weno7z<T, TW>(fw, TW(1.0e-20)));
            pin(o);
#pragma unroll
            for (int j = 1; j < 7; j++) pin(fw[j]);
            out[r] = o;
#pragma unroll
            for (int j = 0; j < 6; j++) fw[j] = fw[j + 1];
        }
    } else if constexpr (SCH == WENO3JS) {
#pragma unroll
        for (int r = 0; r < R; r++) out[r] = TS(weno3js<T, TW>(win + r));
    } else if constexpr (SCH == UPWIND1ST) {
        // src/upwind1st.cpp:56-113: the upwind cell itself
#pragma unroll
        for (int r = 0; r < R; r++) out[r] = TS(win[r + 1]);
    } else if constexpr (SCH == HCCS78) {
        // blend on the inner interfaces ts..nloc-ts (src/hccs78.cpp:56-60),
        // WENO7-Z on the edge interfaces (:207-304)
        T fw[8];
#pragma unroll
        for (int j = 0; j < 7; j++) fw[j] = win[j];
#pragma unroll
        for (int r = 0; r < R; r++) {
            fw[7] = win[r + 7];
            bool edge = ml[r] < z.ts || ml[r] > z.nloc - z.ts;
            TS o = edge ? TS(weno7z<T, TW>(fw, TW(1.0e-20))) : TS(hccs78_inner<T, TW>(fw));
            pin(o);
#pragma unroll
            for (int j = 1; j < 8; j++) pin(fw[j]);
            out[r] = o;
#pragma unroll
            for (int j = 0; j < 7; j++) fw[j] = fw[j + 1];
        }
    } else if constexpr (SCH == CRWENO5) {
        Chunk<R, TS> ch;
        // Rows in order: a rolling window (f1..f4, the differences d1..d4 and
        // the second-difference terms e1..e3 of the current row), where the
        // values the next row reuses and the row's own coefficients pass
        // through an empty asm.  Left free,
        // the compiler evaluates every row's weights before the first
        // elimination step and holds about ten values per row (233 registers,
        // 2 waves per SIMD, against 164 and 3 waves in order).
        T f1 = win[1], f2 = win[2], f3 = win[3];
        T d1 = f1 - win[0], d2 = f2 - f1, d3 = f3 - f2;
        T e1 = T(13.0 / 12.0) * sq(d2 - d1), e2 = T(13.0 / 12.0) * sq(d3 - d2);
#pragma unroll
        for (int r = 0; r < R; r++) {
            T f4 = win[r + 4];
            T d4 = f4 - f3;
            T e3 = T(13.0 / 12.0) * sq(d4 - d3);
            T b0, b1, b2, q0, q1, q2, h2;
            w5_diff<T>(f2, f3, d1, d2, d3, d4, e1, e2, e3, b0, b1, b2, q0, q1, q2, h2);
            TW P0, P1, P2;   // weights with unit ideal weights
            w5_weights<TW>(TW(b0), TW(b1), TW(b2), TW(1), TW(1), TW(1), P0, P1, P2);
            // Every row is scaled by a factor of its own, which leaves the
            // tridiagonal solution unchanged.  The explicit WENO5-JS closure
            // (src/crweno5.cpp:87-129) becomes the row (sum n_k) x = sum n_k q_k
            // with the unnormalised JS weights n_k; the compact row
            // (src/crweno5.cpp:133-162) is scaled by 3 times the weight
            // denominator.
            TW n0 = TW(0.1) * P0, n1 = TW(0.6) * P1, n2 = TW(0.3) * P2;
            TW be = n0 + n1 + n2;
            TW de = n0 * TW(q0) + n1 * TW(q1) + n2 * TW(q2);
            TW m0 = TW(0.2) * P0, m1 = TW(0.5) * P1, m2 = TW(0.3) * P2;
            TW lo = TW(2) * m0 + m1;
            TW di = m0 + TW(2) * (m1 + m2);
            TW hi = m2;
            TW rhs = (m0 * TW(f1) + (TW(5) * (m0 + m1) + m2) * TW(f2) + (m1 + TW(5) * m2) * TW(f3)) * TW(0.5);
            int m = m0r + mstep * r;
            bool valid = full || m < z.nrows;
            bool expl = ml[r] <= z.ts || ml[r] >= z.nloc - z.ts;
            TS a = (valid && !expl) ? TS(lo) : TS(0);
            TS b = valid ? TS(expl ? be : di) : TS(1);
            TS c = (valid && !expl) ? TS(hi) : TS(0);
            TS d = valid ? TS(expl ? de : rhs) : TS(0);
            // the previous row's elimination step is done before this row's
            // coefficients are released, so that rows do not queue up for it
            if (r > 0) {
                pin(ch.Af[r - 1]); pin(ch.Cf[r - 1]); pin(ch.Df[r - 1]);
            }
            asm volatile("" : "+v"(f2), "+v"(f3), "+v"(f4), "+v"(d3), "+v"(d4), "+v"(e2), "+v"(e3),
                              "+v"(a), "+v"(b), "+v"(c), "+v"(d));
            ch.row(r, a, b, c, d);
            f1 = f2; f2 = f3; f3 = f4; d1 = d2; d2 = d3; d3 = d4; e1 = e2; e2 = e3;
        }
        ch.solve(h);
#pragma unroll
        for (int r = 0; r < R; r++) out[r] = ch.Df[r];
    } else if constexpr (SCH == HCCS77) {
        // compact rows on ts+1..nloc-ts-1, WENO7-Z rows at ts and nloc-ts and
        // on the edges (src/hccs77.cpp:86-163, :349-455)
        Chunk<R, TS> ch;
        T fw[7];
#pragma unroll
        for (int j = 0; j < 6; j++) fw[j] = win[j];
#pragma unroll
        for (int r = 0; r < R; r++) {
            fw[6] = win[r + 6];
            const T *f = fw;
            int m = m0r + mstep * r;
            bool valid = full || m < z.nrows;
            bool expl = ml[r] <= z.ts || ml[r] >= z.nloc - z.ts;
            TW a = TW(0), b = TW(1), c = TW(0), d = TW(0);
            if (valid) {
                if (expl) d = weno7z<T, TW>(f, TW(1.0e-20));
                else hccs77_row<T, TW>(f, a, b, c, d);
            }
            TS as = TS(a), bs = TS(b), cs = TS(c), ds = TS(d);
            if (r > 0) {
                pin(ch.Af[r - 1]); pin(ch.Cf[r - 1]); pin(ch.Df[r - 1]);
            }
            pin(as); pin(bs); pin(cs); pin(ds);
#pragma unroll
            for (int j = 1; j < 7; j++) pin(fw[j]);
            ch.row(r, as, bs, cs, ds);
#pragma unroll
            for (int j = 0; j < 6; j++) fw[j] = fw[j + 1];
        }
        ch.solve(h);
#pragma unroll
        for (int r = 0; r < R; r++) out[r] = ch.Df[r];
    } else {  // HCCS
        Chunk<R, TS> ch;
        T fw[7];   // 7-point window for the closures, 5-point from fw+1
#pragma unroll
        for (int j = 0; j < 6; j++) fw[j] = win[j];
#pragma unroll
        for (int r = 0; r < R; r++) {
            fw[6] = win[r + 6];
            const T *f = fw;
            int m = m0r + mstep * r;
            bool valid = full || m < z.nrows;
            bool expl = ml[r] <= z.ts || ml[r] >= z.nloc - z.ts;
            TW a = TW(0), b = TW(1), c = TW(0), d = TW(0);
            if (valid) {
                if (expl) {
                    d = weno7z<T, TW>(f, TW(1.0e-40));
                } else {
                    hccs_row<T, TW>(f[1], f[2], f[3], f[4], f[5], a, b, c, d);
                }
            }
            TS as = TS(a), bs = TS(b), cs = TS(c), ds = TS(d);
            if (r > 0) {
                pin(ch.Af[r - 1]); pin(ch.Cf[r - 1]); pin(ch.Df[r - 1]);
            }
            pin(as); pin(bs); pin(cs); pin(ds);
#pragma unroll
            for (int j = 1; j < 7; j++) pin(fw[j]);
            ch.row(r, as, bs, cs, ds);
#pragma unroll
            for (int j = 0; j < 6; j++) fw[j] = fw[j + 1];
        }
        ch.solve(h);
#pragma unroll
        for (int r = 0; r < R; r++) out[r] = ch.Df[r];
    }
}

// First derivative at cell c of a line buffer L (cd2 / cd6 / cd8); `inner` selects
// the inner-loop form of cd8 (src/centraldifference6th.cpp:42-48,
// src/centraldifference8th.cpp:52-66 inner, :127-131 edges).
template <typename T>
__device__ __forceinline__ T central_diff(const T *L, int c, bool inner, int vis, T inv_h)
{
    if (vis == VIS_CD2) return T(0.5) * (L[c + 1] - L[c - 1]) * inv_h;   // src/centraldifference2nd.cpp:38-40
    if (vis == VIS_CD6) {
        return (T(1.0 / 60.0) * (L[c + 3] - L[c - 3]) + T(-3.0 / 20.0) * (L[c + 2] - L[c - 2])
                + T(3.0 / 4.0) * (L[c + 1] - L[c - 1])) * inv_h;
    }
    if (inner) {
        T an = T(-1.0 / 280.0) * L[c - 4] + T(4.0 / 105.0) * L[c - 3] + T(-1.0 / 5.0) * L[c - 2] + T(4.0 / 5.0) * L[c - 1];
        T ap = T(4.0 / 5.0) * L[c + 1] + T(-1.0 / 5.0) * L[c + 2] + T(4.0 / 105.0) * L[c + 3] + T(-1.0 / 280.0) * L[c + 4];
        return (ap - an) * inv_h;
    }
    return (T(-1.0 / 280.0) * (L[c + 4] - L[c - 4]) + T(4.0 / 105.0) * (L[c + 3] - L[c - 3])
            + T(-1.0 / 5.0) * (L[c + 2] - L[c - 2]) + T(4.0 / 5.0) * (L[c + 1] - L[c - 1])) * inv_h;
}

}  // namespace nucdev
