// nuc_lines.h -- the per-line interface computation shared by the production
// sweep kernels and the per-Field operator kernels.
//
// One wave owns one line held in LDS on [-HALO, n+HALO).  Lane k owns the
// interfaces m = k*CPT + r.  Rank-local zones follow the reference's
// inner/boundary split per MPI block of length nloc (src/fieldOperator.cpp:387-465):
//   explicit boundary interfaces  ml in [0, ts-1] U [nloc-ts+1, nloc]
//   inner interfaces              ml in [ts, nloc-ts]
//   compact schemes: rows ml = ts and ml = nloc-ts are explicit closures,
//   ts < ml < nloc-ts are tridiagonal rows (src/crweno5.cpp:83-206).
#pragma once
#include "nuc_device.h"

namespace nucdev {

constexpr int HALO = 4;

struct Zone {
    int nrows;  // interfaces computed on this line (n periodic, n+1 with ghosts)
    int nloc;   // rank-local line length
    int ts;     // inviscid tilesize of the scheme
};

template <int SCH, typename T>
__device__ __forceinline__ T explicit_pos(const T *P, int m, bool inner)
{
    if constexpr (SCH == WENO5JS || SCH == CRWENO5) {
        return weno5js<T>(P[m - 3], P[m - 2], P[m - 1], P[m], P[m + 1]);
    } else if constexpr (SCH == WENO5Z) {
        return weno5z<T>(P[m - 3], P[m - 2], P[m - 1], P[m], P[m + 1], inner);
    } else {
        T f[7] = {P[m - 4], P[m - 3], P[m - 2], P[m - 1], P[m], P[m + 1], P[m + 2]};
        return weno7z<T>(f, SCH == WENO7JS ? T(1.0e-20) : T(1.0e-40));
    }
}

template <int SCH, typename T>
__device__ __forceinline__ T explicit_neg(const T *M, int m)
{
    if constexpr (SCH == WENO5JS || SCH == CRWENO5) {
        return weno5js<T>(M[m + 2], M[m + 1], M[m], M[m - 1], M[m - 2]);
    } else if constexpr (SCH == WENO5Z) {
        return weno5z<T>(M[m + 2], M[m + 1], M[m], M[m - 1], M[m - 2], false);
    } else {
        T f[7] = {M[m + 3], M[m + 2], M[m + 1], M[m], M[m - 1], M[m - 2], M[m - 3]};
        return weno7z<T>(f, SCH == WENO7JS ? T(1.0e-20) : T(1.0e-40));
    }
}

// Reconstructed f^+ (into Fp) and f^- (into Fm) at the lane's interfaces.
// zone_mask: bit0 = inner interfaces, bit1 = boundary interfaces (used by the
// operator hooks; the sweeps pass 3).  Interfaces not selected are left
// untouched in Fp/Fm.
template <int SCH, int CPT, typename T>
__device__ __forceinline__ void line_interfaces(const T *P, const T *M, const Zone &z, int lane,
                                                T Fp[CPT], T Fm[CPT])
{
    if constexpr (SCH == CRWENO5 || SCH == HCCS) {
        T ap[CPT], bp[CPT], cp[CPT], dp[CPT];
        T am[CPT], bm[CPT], cm[CPT], dm[CPT];
#pragma unroll
        for (int r = 0; r < CPT; r++) {
            int m = lane * CPT + r;
            ap[r] = T(0); bp[r] = T(1); cp[r] = T(0); dp[r] = T(0);
            am[r] = T(0); bm[r] = T(1); cm[r] = T(0); dm[r] = T(0);
            if (m < z.nrows) {
                int ml = m % z.nloc;
                if (ml <= z.ts || ml >= z.nloc - z.ts) {
                    dp[r] = explicit_pos<SCH, T>(P, m, true);
                    dm[r] = explicit_neg<SCH, T>(M, m);
                } else {
                    T lo, di, hi, rhs;
                    if constexpr (SCH == CRWENO5)
                        crweno5_row<T>(P[m - 3], P[m - 2], P[m - 1], P[m], P[m + 1], lo, di, hi, rhs);
                    else
                        hccs_row<T>(P[m - 3], P[m - 2], P[m - 1], P[m], P[m + 1], lo, di, hi, rhs);
                    ap[r] = lo; bp[r] = di; cp[r] = hi; dp[r] = rhs;
                    if constexpr (SCH == CRWENO5)
                        crweno5_row<T>(M[m + 2], M[m + 1], M[m], M[m - 1], M[m - 2], lo, di, hi, rhs);
                    else
                        hccs_row<T>(M[m + 2], M[m + 1], M[m], M[m - 1], M[m - 2], lo, di, hi, rhs);
                    // negative sweep: roles of sub- and super-diagonal swap (src/crweno5.cpp:329-332)
                    am[r] = hi; bm[r] = di; cm[r] = lo; dm[r] = rhs;
                }
            }
        }
        tri_solve_wave<CPT, T>(ap, bp, cp, dp, lane);
        tri_solve_wave<CPT, T>(am, bm, cm, dm, lane);
#pragma unroll
        for (int r = 0; r < CPT; r++) { Fp[r] = dp[r]; Fm[r] = dm[r]; }
    } else {
#pragma unroll
        for (int r = 0; r < CPT; r++) {
            int m = lane * CPT + r;
            Fp[r] = T(0); Fm[r] = T(0);
            if (m < z.nrows) {
                int ml = m % z.nloc;
                bool inner = (ml >= z.ts && ml <= z.nloc - z.ts);
                Fp[r] = explicit_pos<SCH, T>(P, m, inner);
                Fm[r] = explicit_neg<SCH, T>(M, m);
            }
        }
    }
}

// First derivative at cell c of a line buffer L (cd6 / cd8), zone by the
// rank-local position cl (src/centraldifference6th.cpp:42-48,
// src/centraldifference8th.cpp:52-66 inner, :127-131 edges).
template <typename T>
__device__ __forceinline__ T central_diff(const T *L, int c, int cl, int nloc, int tsv, int vis, T inv_h)
{
    if (vis == VIS_CD6) {
        return (T(1.0 / 60.0) * (L[c + 3] - L[c - 3]) + T(-3.0 / 20.0) * (L[c + 2] - L[c - 2])
                + T(3.0 / 4.0) * (L[c + 1] - L[c - 1])) * inv_h;
    }
    if (cl >= tsv && cl < nloc - tsv) {
        T an = T(-1.0 / 280.0) * L[c - 4] + T(4.0 / 105.0) * L[c - 3] + T(-1.0 / 5.0) * L[c - 2] + T(4.0 / 5.0) * L[c - 1];
        T ap = T(4.0 / 5.0) * L[c + 1] + T(-1.0 / 5.0) * L[c + 2] + T(4.0 / 105.0) * L[c + 3] + T(-1.0 / 280.0) * L[c + 4];
        return (ap - an) * inv_h;
    }
    return (T(-1.0 / 280.0) * (L[c + 4] - L[c - 4]) + T(4.0 / 105.0) * (L[c + 3] - L[c - 3])
            + T(-1.0 / 5.0) * (L[c + 2] - L[c - 2]) + T(4.0 / 5.0) * (L[c + 1] - L[c - 1])) * inv_h;
}

}  // namespace nucdev
