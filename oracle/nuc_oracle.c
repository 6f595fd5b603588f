/*
 * nuc_oracle.c -- CPU restatement of NUC3d's per-stage right-hand side and
 * TVD-RK3 advance.  TEST INFRASTRUCTURE ONLY (see nuc_oracle.h): it is the
 * checker for the HIP path and the timed CPU baseline, never the product.
 *
 * Every formula keeps the evaluation order of the reference source line it
 * cites, so that compiled with -ffp-contract=off it reproduces the
 * reference's own scheme objects bit for bit (checked in tests/ against
 * oracle/_ref, the reference's scheme sources compiled unmodified).
 *
 * Rank emulation: the reference runs npx*npy*npz MPI ranks on a periodic
 * Cartesian grid (src/MPICommunicator.cpp:26-56); each rank sweeps its local
 * field with ghost strips received from its neighbours
 * (src/bufferData.cpp:56-192).  Here a global line along the sweep direction is
 * copied once into a periodically extended buffer; the segment of rank r is
 * then exactly the rank-local line with its ghost cells, and the reference's
 * per-rank inner/boundary code is applied to it.
 *
 * ORC_PLANT_* macros: compiled in only by the tests that plant a known
 * mistake (a dropped term, a swapped factor) into a throwaway copy of this
 * library, to show that the closed-form pins of those functions fail on it.
 * The normal build defines none of them.
 */
#include "nuc_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* coefficient tables: src/schemes.hpp:14-42 */
static const double A5[3][3] = {
    {1.0 / 3.0, -7.0 / 6.0, 11.0 / 6.0},
    {-1.0 / 6.0, 5.0 / 6.0, 1.0 / 3.0},
    {1.0 / 3.0, 5.0 / 6.0, -1.0 / 6.0}};
static const double A7[4][4] = {
    {-3.0 / 12.0, 13.0 / 12.0, -23.0 / 12.0, 25.0 / 12.0},
    {1.0 / 12.0, -5.0 / 12.0, 13.0 / 12.0, 3.0 / 12.0},
    {-1.0 / 12.0, 7.0 / 12.0, 7.0 / 12.0, -1.0 / 12.0},
    {3.0 / 12.0, 13.0 / 12.0, -5.0 / 12.0, 1.0 / 12.0}};
static const double C5[3] = {0.1, 0.6, 0.3};
static const double CCR[3] = {0.2, 0.5, 0.3};
static const double CUC[3] = {2.0 / 7.0, 4.0 / 7.0, 1.0 / 7.0};
static const double C7[4] = {1.0 / 35.0, 12.0 / 35.0, 18.0 / 35.0, 4.0 / 35.0};
static const double G0 = 13.0 / 12.0;
static const double G1 = 1.0 / 4.0;
static const double CD8[4] = {4.0 / 5.0, -1.0 / 5.0, 4.0 / 105.0, -1.0 / 280.0};
static const double CD6[3] = {3.0 / 4.0, -3.0 / 20.0, 1.0 / 60.0};

/* epsilons: src/weno5js.cpp:8, src/weno5z.cpp:7, src/weno7js.cpp:14,
 * src/crweno5.cpp:14, src/hccs.cpp:14 */
#define EPS_JS 1.0e-6
#define EPS_Z 1.0e-40
#define EPS_7 1.0e-20

/* The schemes raise to their integer member p (= 2) through std::pow at run
 * time (e.g. src/weno5js.cpp:97-99, src/weno5z.cpp:74-100), which calls libm
 * pow(x, 2.0); a literal std::pow(x, 2) folds to x*x.  The two can differ in
 * the last bit, so the member-p sites go through libm here as well. */
static volatile double P_MEMBER = 2.0;
#define POWP(x) pow((x), P_MEMBER)

/* bufferSize per scheme, src/fieldOperator.cpp:128-181 */
int orc_scheme_halo(int scheme)
{
    if (scheme == ORC_WENO7JS || scheme == ORC_HCCS || scheme == ORC_HCCS78 || scheme == ORC_HCCS77) return 4;
    if (scheme == ORC_WENO3JS) return 2;
    if (scheme == ORC_UPWIND1ST) return 1;
    return 3;
}

/* cells on each side a viscous scheme reads: cd2 1, cd6 3, cd8 4 */
static int vis_reach(int vis)
{
    return vis == ORC_CD8 ? 4 : (vis == ORC_CD2 ? 1 : 3);
}

/* ---------------------------------------------------------------------------
 * Point formulas.  f[] is the stencil in "upwind coordinates": for the
 * positive (f+) sweep f[2+z] = cell(m-1+z), for the negative sweep
 * f[2+z] = cell(m-z), m the interface index.  Both orientations use the same
 * expression (src/weno5js.cpp:62-111 vs :143-190).
 * ------------------------------------------------------------------------- */

/* WENO5-JS, src/weno5js.cpp:62-111 */
static double weno5js_pt(const double *f)
{
    double is0 = G0 * pow((f[0] - 2.0 * f[1] + f[2]), 2) + G1 * pow((f[0] - 4.0 * f[1] + 3.0 * f[2]), 2);
    double is1 = G0 * pow((f[1] - 2.0 * f[2] + f[3]), 2) + G1 * pow((f[1] - f[3]), 2);
    double is2 = G0 * pow((f[2] - 2.0 * f[3] + f[4]), 2) + G1 * pow((3.0 * f[2] - 4.0 * f[3] + f[4]), 2);
    double q30 = A5[0][0] * f[0] + A5[0][1] * f[1] + A5[0][2] * f[2];
    double q31 = A5[1][0] * f[1] + A5[1][1] * f[2] + A5[1][2] * f[3];
    double q32 = A5[2][0] * f[2] + A5[2][1] * f[3] + A5[2][2] * f[4];
    double alpha0 = C5[0] / POWP((EPS_JS + is0));
    double alpha1 = C5[1] / POWP((EPS_JS + is1));
    double alpha2 = C5[2] / POWP((EPS_JS + is2));
    double alphaSum = alpha0 + alpha1 + alpha2;
    double omega0 = alpha0 / alphaSum;
    double omega1 = alpha1 / alphaSum;
    double omega2 = alpha2 / alphaSum;
    return omega0 * q30 + omega1 * q31 + omega2 * q32;
}

/* WENO5-Z, src/weno5z.cpp:72-110.  typo!=0 reproduces the positive inner
 * sweep's is2, whose first square reads f[i+1] twice (src/weno5z.cpp:80). */
static double weno5z_pt(const double *f, int typo)
{
    double is0 = G0 * POWP((f[0] - 2.0 * f[1] + f[2])) + G1 * POWP((f[0] - 4.0 * f[1] + 3.0 * f[2]));
    double is1 = G0 * POWP((f[1] - 2.0 * f[2] + f[3])) + G1 * POWP((f[1] - f[3]));
    double is2 = typo ? G0 * POWP((f[2] - 2.0 * f[3] + f[3])) + G1 * POWP((3.0 * f[2] - 4.0 * f[3] + f[4]))
                      : G0 * POWP((f[2] - 2.0 * f[3] + f[4])) + G1 * POWP((3.0 * f[2] - 4.0 * f[3] + f[4]));
    double q30 = A5[0][0] * f[0] + A5[0][1] * f[1] + A5[0][2] * f[2];
    double q31 = A5[1][0] * f[1] + A5[1][1] * f[2] + A5[1][2] * f[3];
    double q32 = A5[2][0] * f[2] + A5[2][1] * f[3] + A5[2][2] * f[4];
    double tau = fabs(is2 - is0);
    double alpha0 = C5[0] * (1.0 + POWP(tau / (is0 + EPS_Z)));
    double alpha1 = C5[1] * (1.0 + POWP(tau / (is1 + EPS_Z)));
    double alpha2 = C5[2] * (1.0 + POWP(tau / (is2 + EPS_Z)));
    double alphaSum = alpha0 + alpha1 + alpha2;
    double omega0 = alpha0 / alphaSum;
    double omega1 = alpha1 / alphaSum;
    double omega2 = alpha2 / alphaSum;
    return omega0 * q30 + omega1 * q31 + omega2 * q32;
}

/* 7-point WENO with Z weights (src/weno7js.cpp:79-133; hccs closures use the
 * same formula with eps=1e-40, src/hccs.cpp:85-135; hccs78 edges with 1e-20,
 * src/hccs78.cpp:246-304).  f[3+z], z=-3..3.  *asum receives alphaSum. */
static double weno7z_core(const double *f, double eps, int member_p, double *asum)
{
    double is0 = f[0] * (547.0 * f[0] - 3882.0 * f[1] + 4642.0 * f[2] - 1854.0 * f[3])
               + f[1] * (7043.0 * f[1] - 17246.0 * f[2] + 7042.0 * f[3])
               + f[2] * (11003.0 * f[2] - 9402.0 * f[3])
               + 2107.0 * f[3] * f[3];
    double is1 = f[1] * (267.0 * f[1] - 1642.0 * f[2] + 1602.0 * f[3] - 494.0 * f[4])
               + f[2] * (2843.0 * f[2] - 5966.0 * f[3] + 1922.0 * f[4])
               + f[3] * (3443.0 * f[3] - 2522.0 * f[4])
               + 547.0 * f[4] * f[4];
    double is2 = f[2] * (547.0 * f[2] - 2522.0 * f[3] + 1922.0 * f[4] - 494.0 * f[5])
               + f[3] * (3443.0 * f[3] - 5966.0 * f[4] + 1602.0 * f[5])
               + f[4] * (2843.0 * f[4] - 1642.0 * f[5])
               + 267.0 * f[5] * f[5];
    double is3 = f[3] * (2107.0 * f[3] - 9402.0 * f[4] + 7042.0 * f[5] - 1854.0 * f[6])
               + f[4] * (11003.0 * f[4] - 17246.0 * f[5] + 4642.0 * f[6])
               + f[5] * (7043.0 * f[5] - 3882.0 * f[6])
               + 547.0 * f[6] * f[6];
    double q40 = A7[0][0] * f[0] + A7[0][1] * f[1] + A7[0][2] * f[2] + A7[0][3] * f[3];
    double q41 = A7[1][0] * f[1] + A7[1][1] * f[2] + A7[1][2] * f[3] + A7[1][3] * f[4];
    double q42 = A7[2][0] * f[2] + A7[2][1] * f[3] + A7[2][2] * f[4] + A7[2][3] * f[5];
    double q43 = A7[3][0] * f[3] + A7[3][1] * f[4] + A7[3][2] * f[5] + A7[3][3] * f[6];
    double tau7 = fabs(is0 - is3);
    double alpha0 = C7[0] * (1.0 + (member_p ? POWP(tau7 / (eps + is0)) : pow(tau7 / (eps + is0), 2)));
    double alpha1 = C7[1] * (1.0 + (member_p ? POWP(tau7 / (eps + is1)) : pow(tau7 / (eps + is1), 2)));
    double alpha2 = C7[2] * (1.0 + (member_p ? POWP(tau7 / (eps + is2)) : pow(tau7 / (eps + is2), 2)));
    double alpha3 = C7[3] * (1.0 + (member_p ? POWP(tau7 / (eps + is3)) : pow(tau7 / (eps + is3), 2)));
    double alphaSum = alpha0 + alpha1 + alpha2 + alpha3;
    double omega0 = alpha0 / alphaSum;
    double omega1 = alpha1 / alphaSum;
    double omega2 = alpha2 / alphaSum;
    double omega3 = alpha3 / alphaSum;
    *asum = alphaSum;
    return omega0 * q40 + omega1 * q41 + omega2 * q42 + omega3 * q43;
}

static double weno7z_pt(const double *f, double eps, int member_p)
{
    double asum;
    return weno7z_core(f, eps, member_p, &asum);
}

/* hccs78 inner point (src/hccs78.cpp:66-117): the WENO7-Z value blended with
 * the 8-point central value q8 through theta = 1/(1 + (alphaSum-1)^2).
 * f[3+z], z=-3..4 (one downwind point more than WENO7). */
static double hccs78_pt(const double *f)
{
    double q8 = (-3.0 * (f[7] + f[0]) + 29.0 * (f[6] + f[1]) - 139.0 * (f[5] + f[2]) + 533.0 * (f[4] + f[3])) / 840.0;
    double alphaSum;
    double w = weno7z_core(f, EPS_7, 1, &alphaSum);
    double theta = 1.0 / (1.0 + POWP(alphaSum - 1.0));
    return (1.0 - theta) * w + theta * q8;
}

/* hccs77 interior row, src/hccs77.cpp:86-163 (positive; the negative sweep
 * :227-305 swaps the roles of a and c): WENO7-Z (eps 1e-20) and the 7th-order
 * upwind compact row blended through theta = 1/(1 + (alphaSum-1)^2); theta >
 * 0.5 takes the compact row alone.  f[3+z], z=-3..3. */
static void hccs77_row(const double *f, double *lo, double *di, double *hi, double *rhs)
{
    double alphaSum;
    double w = weno7z_core(f, EPS_7, 1, &alphaSum);
    double q7 = (-f[1] + 19.0 * f[2] + 239.0 * f[3] + 159.0 * f[4] + 4.0 * f[5]) / 420.0;
    double theta = 1.0 / (1.0 + POWP(alphaSum - 1.0));
    if (theta > 0.5) {
        *lo = 2.0 / 7.0; *di = 4.0 / 7.0; *hi = 1.0 / 7.0; *rhs = q7;
    } else {
        *lo = CUC[0] * theta;
        *di = 1.0 + (CUC[1] - 1.0) * theta;
        *hi = CUC[2] * theta;
        *rhs = w * (1.0 - theta) + q7 * theta;
    }
}

/* WENO3-JS, src/weno3js.cpp:80-93: f[1+z], z=-1..1 */
static double weno3js_pt(const double *f)
{
    double is0 = pow((f[0] - f[1]), 2);
    double is1 = pow((f[1] - f[2]), 2);
    double q20 = -0.5 * f[0] + 1.5 * f[1];
    double q21 = 0.5 * f[1] + 0.5 * f[2];
    double aa0 = 1.0 / 3.0 / POWP(EPS_JS + is0);
    double aa1 = 2.0 / 3.0 / POWP(EPS_JS + is1);
    double w0 = aa0 / (aa0 + aa1);
    double w1 = aa1 / (aa0 + aa1);
    return w0 * q20 + w1 * q21;
}

/* CRWENO5 interior row, upwind coordinates (src/crweno5.cpp:133-162): the
 * (lo, di, hi) triple multiplies the interfaces (behind, here, ahead) in the
 * upwind direction. */
static void crweno5_row(const double *f, double *lo, double *di, double *hi, double *rhs)
{
    double is0 = G0 * pow((f[0] - 2.0 * f[1] + f[2]), 2) + G1 * pow((f[0] - 4.0 * f[1] + 3.0 * f[2]), 2);
    double is1 = G0 * pow((f[1] - 2.0 * f[2] + f[3]), 2) + G1 * pow((f[1] - f[3]), 2);
    double is2 = G0 * pow((f[2] - 2.0 * f[3] + f[4]), 2) + G1 * pow((3.0 * f[2] - 4.0 * f[3] + f[4]), 2);
    double alpha0 = CCR[0] / POWP((EPS_JS + is0));
    double alpha1 = CCR[1] / POWP((EPS_JS + is1));
    double alpha2 = CCR[2] / POWP((EPS_JS + is2));
    double alphaSum = alpha0 + alpha1 + alpha2;
    double omega0 = alpha0 / alphaSum;
    double omega1 = alpha1 / alphaSum;
    double omega2 = alpha2 / alphaSum;
    *lo = 2.0 / 3.0 * omega0 + 1.0 / 3.0 * omega1;
    *di = 1.0 / 3.0 * omega0 + 2.0 / 3.0 * (omega1 + omega2);
    *hi = 1.0 / 3.0 * omega2;
    *rhs = 1.0 / 6.0 * omega0 * f[1] + (5.0 * (omega0 + omega1) + omega2) * f[2] / 6.0
         + (omega1 + 5.0 * omega2) * f[3] / 6.0;
}

/* hccs interior row (src/hccs.cpp:138-183): Z-type CRWENO weights blended
 * with the 7th-order upwind compact scheme through theta. */
static void hccs_row(const double *f, double *lo, double *di, double *hi, double *rhs)
{
    double is0 = G0 * pow((f[0] - 2.0 * f[1] + f[2]), 2) + G1 * pow((f[0] - 4.0 * f[1] + 3.0 * f[2]), 2);
    double is1 = G0 * pow((f[1] - 2.0 * f[2] + f[3]), 2) + G1 * pow((f[1] - f[3]), 2);
    double is2 = G0 * pow((f[2] - 2.0 * f[3] + f[4]), 2) + G1 * pow((3.0 * f[2] - 4.0 * f[3] + f[4]), 2);
    double q7 = (-f[0] + 19.0 * f[1] + 239.0 * f[2] + 159.0 * f[3] + 4.0 * f[4]) / 420.0;
    double tau = fabs(is2 - is0);
    double alpha0 = CCR[0] * (1.0 + pow(tau / (is0 + EPS_Z), 2));
    double alpha1 = CCR[1] * (1.0 + pow(tau / (is1 + EPS_Z), 2));
    double alpha2 = CCR[2] * (1.0 + pow(tau / (is2 + EPS_Z), 2));
    double alphaSum = alpha0 + alpha1 + alpha2;
    double omega0 = alpha0 / alphaSum;
    double omega1 = alpha1 / alphaSum;
    double omega2 = alpha2 / alphaSum;
    double theta = 1.0 / (1.0 + pow(alphaSum - 1.0, 2));
    double blend = (1.0 - theta) / theta;
#if defined(ORC_PLANT_HCCS_SWAP)
    blend = theta / (1.0 - theta);
#elif defined(ORC_PLANT_HCCS_INV)
    blend = 1.0 / theta;
#endif
    if (theta > 0.5) {
        *lo = 2.0 / 7.0;
        *di = 4.0 / 7.0;
        *hi = 1.0 / 7.0;
        *rhs = q7;
    } else {
        *lo = (2.0 / 3.0 * omega0 + 1.0 / 3.0 * omega1) * blend + CUC[0];
        *di = (1.0 / 3.0 * omega0 + 2.0 / 3.0 * (omega1 + omega2)) * blend + CUC[1];
        *hi = (1.0 / 3.0 * omega2) * blend + CUC[2];
        *rhs = (1.0 / 6.0 * omega0 * f[1] + (5.0 * (omega0 + omega1) + omega2) * f[2] / 6.0
                + (omega1 + 5.0 * omega2) * f[3] / 6.0) * blend + q7;
    }
}

/* One interior hccs row from its 5-point upwind stencil f[0..4] (cells i-2..i+2
 * of src/hccs.cpp:138-183): out = {lo, di, hi, rhs}. */
void orc_hccs_row(const double *f, double *out) { hccs_row(f, out, out + 1, out + 2, out + 3); }

/* gather a stencil z = -r .. r+extra for interface m in upwind coordinates */
static inline void gather_ext(const double *f, int m, int upwind, int r, int extra, double *st)
{
    if (upwind > 0)
        for (int z = -r; z <= r + extra; z++) st[r + z] = f[m - 1 + z];
    else
        for (int z = -r; z <= r + extra; z++) st[r + z] = f[m - z];
}

static inline void gather(const double *f, int m, int upwind, int r, double *st)
{
    gather_ext(f, m, upwind, r, 0, st);
}

static double explicit_pt(int scheme, int upwind, const double *f, int m, int inner)
{
    double st[8];
    switch (scheme) {
    case ORC_WENO3JS:
        gather(f, m, upwind, 1, st);
        return weno3js_pt(st);
    case ORC_UPWIND1ST:
        /* src/upwind1st.cpp:56-74 (+: cell m-1), :95-113 (-: cell m) */
        gather(f, m, upwind, 1, st);
        return st[1];
    case ORC_HCCS77:
        /* edges (src/hccs77.cpp:349-455) and the explicit first/last compact
         * rows (:150-157): WENO7-Z with eps 1e-20 */
        gather(f, m, upwind, 3, st);
        return weno7z_pt(st, EPS_7, 1);
    case ORC_HCCS78:
        /* inner: blend (src/hccs78.cpp:47-117); edges: WENO7-Z (:207-304) */
        if (!inner) {
            gather(f, m, upwind, 3, st);
            return weno7z_pt(st, EPS_7, 1);
        }
        gather_ext(f, m, upwind, 3, 1, st);
        return hccs78_pt(st);
    case ORC_WENO5JS:
    case ORC_CRWENO5:
        gather(f, m, upwind, 2, st);
        return weno5js_pt(st);
    case ORC_WENO5Z:
        gather(f, m, upwind, 2, st);
        return weno5z_pt(st, inner && upwind > 0);
    case ORC_WENO7JS:
        gather(f, m, upwind, 3, st);
        return weno7z_pt(st, EPS_7, 1);
    default: /* ORC_HCCS closures */
        gather(f, m, upwind, 3, st);
        return weno7z_pt(st, EPS_Z, 0);
    }
}

/* Non-periodic Thomas solve on the rank-local compact rows, as written in
 * src/crweno5.cpp:208-219 (identical in src/hccs.cpp:246-257). */
static void thomas(double *a, double *b, double *c, double *d, int M, double *x)
{
    for (int i = 1; i < M; i++) {
        c[i] = c[i] / (b[i] - c[i - 1] * a[i]);
        d[i] = (d[i] - d[i - 1] * a[i]) / (b[i] - c[i - 1] * a[i]);
    }
    x[M - 1] = d[M - 1];
    for (int i = M - 2; i >= 0; i--) x[i] = d[i] - c[i] * x[i + 1];
}

void orc_line_reconstruct(int scheme, int upwind, const double *f, int n, int ts, double *out)
{
    /* boundary interfaces: interpolationBoundaryL/R (src/weno5js.cpp:216-628),
     * out[0..ts-1] and out[n-ts+1..n] */
    for (int m = 0; m < ts; m++) out[m] = explicit_pt(scheme, upwind, f, m, 0);
    for (int m = n - ts + 1; m <= n; m++) out[m] = explicit_pt(scheme, upwind, f, m, 0);
    if (scheme != ORC_CRWENO5 && scheme != ORC_HCCS && scheme != ORC_HCCS77) {
        /* inner sweep: i = ts-1 .. n-ts-1 writes out[i+1] (src/weno5js.cpp:56-111) */
        for (int m = ts; m <= n - ts; m++) out[m] = explicit_pt(scheme, upwind, f, m, 1);
        return;
    }
    /* compact inner sweep: rows r = 0..M-1 <-> interfaces ts+r; the first and
     * last rows are explicit (src/crweno5.cpp:83-129,164-206; src/hccs.cpp:83-136,191-244) */
    int M = n - 2 * ts + 1;
    double *a = (double *)malloc(sizeof(double) * 5 * (size_t)M);
    double *b = a + M, *c = b + M, *d = c + M, *x = d + M;
    for (int r = 0; r < M; r++) {
        int m = ts + r;
        if (r == 0 || r == M - 1) {
            a[r] = 0.0; b[r] = 1.0; c[r] = 0.0;
            d[r] = explicit_pt(scheme, upwind, f, m, 1);
        } else {
            double st[7], lo, di, hi, rhs;
            if (scheme == ORC_HCCS77) {
                gather(f, m, upwind, 3, st);
                hccs77_row(st, &lo, &di, &hi, &rhs);
            } else {
                gather(f, m, upwind, 2, st);
                if (scheme == ORC_CRWENO5) crweno5_row(st, &lo, &di, &hi, &rhs);
                else hccs_row(st, &lo, &di, &hi, &rhs);
            }
            /* positive: lo multiplies interface m-1 (sub-diagonal); negative:
             * the roles of a and c swap (src/crweno5.cpp:329-332) */
            if (upwind > 0) { a[r] = lo; b[r] = di; c[r] = hi; }
            else { c[r] = lo; b[r] = di; a[r] = hi; }
            d[r] = rhs;
        }
    }
    thomas(a, b, c, d, M, x);
    for (int r = 0; r < M; r++) out[ts + r] = x[r];
    free(a);
}

void orc_line_diff(int scheme, const double *f, int n, int ts, double h, double *out)
{
    if (scheme == ORC_CD2) {
        /* src/centraldifference2nd.cpp:38-40 (inner) and :75-78, :114-117 (edges) */
        for (int i = 0; i < n; i++) out[i] = 0.5 * (f[i + 1] - f[i - 1]) / h;
        return;
    }
    if (scheme == ORC_CD6) {
        /* src/centraldifference6th.cpp:42-48 (inner) and :113-115,175-177 (edges): same expression */
        for (int i = 0; i < n; i++)
            out[i] = (CD6[2] * (f[i + 3] - f[i - 3]) + CD6[1] * (f[i + 2] - f[i - 2]) + CD6[0] * (f[i + 1] - f[i - 1])) / h;
        return;
    }
    /* cd8: inner i in [ts, n-ts) (src/centraldifference8th.cpp:52-66), edges
     * from the ghost-aware form (:127-131, :196-200) */
    for (int i = 0; i < n; i++) {
        if (i >= ts && i < n - ts) {
            double an = (CD8[3] * f[i - 4] + CD8[2] * f[i - 3] + CD8[1] * f[i - 2] + CD8[0] * f[i - 1]);
            double ap = (CD8[0] * f[i + 1] + CD8[1] * f[i + 2] + CD8[2] * f[i + 3] + CD8[3] * f[i + 4]);
            out[i] = (ap - an) / h;
        } else {
            out[i] = (CD8[3] * (f[i + 4] - f[i - 4]) + CD8[2] * (f[i + 3] - f[i - 3])
                      + CD8[1] * (f[i + 2] - f[i - 2]) + CD8[0] * (f[i + 1] - f[i - 1])) / h;
        }
    }
}

/* ---------------------------------------------------------------------------
 * Field-shaped wrappers (reference calling convention, used against oracle/_ref)
 * ------------------------------------------------------------------------- */
static void build_ext(const double *in, int n1, int n2, int j, int k, const double *gL, const double *gR,
                      int bf, double *ext /* n1+2bf */)
{
    for (int g = 0; g < bf; g++) {
        ext[g] = gL[(size_t)bf * n2 * k + (size_t)bf * j + g];
        ext[bf + n1 + g] = gR[(size_t)bf * n2 * k + (size_t)bf * j + g];
    }
    for (int i = 0; i < n1; i++) ext[bf + i] = in[(size_t)n1 * n2 * k + (size_t)n1 * j + i];
}

void orc_field_reconstruct(int scheme, int upwind, int ts, const double *in, int n1, int n2, int n3,
                           const double *gL, const double *gR, int bf, double *out)
{
    double *ext = (double *)malloc(sizeof(double) * (n1 + 2 * bf));
    for (int k = 0; k < n3; k++)
        for (int j = 0; j < n2; j++) {
            build_ext(in, n1, n2, j, k, gL, gR, bf, ext);
            orc_line_reconstruct(scheme, upwind, ext + bf, n1, ts,
                                 out + (size_t)(n1 + 1) * n2 * k + (size_t)(n1 + 1) * j);
        }
    free(ext);
}

void orc_field_diff(int scheme, int ts, const double *in, double h, int n1, int n2, int n3,
                    const double *gL, const double *gR, int bf, double *out)
{
    double *ext = (double *)malloc(sizeof(double) * (n1 + 2 * bf));
    for (int k = 0; k < n3; k++)
        for (int j = 0; j < n2; j++) {
            build_ext(in, n1, n2, j, k, gL, gR, bf, ext);
            orc_line_diff(scheme, ext + bf, n1, ts, h, out + (size_t)n1 * n2 * k + (size_t)n1 * j);
        }
    free(ext);
}

/* TVD-RK3 Shu-Osher stages, src/TVD-RK3.h:9-12 and src/TVD-RK3.cpp:37-157 */
void orc_rk3(int stage, const double *rhs_in, const double *un, double *ui, double dt, long n)
{
    static const double al[3][2] = {{1.0, 1.0}, {3.0 / 4.0, 1.0 / 4.0}, {1.0 / 3.0, 2.0 / 3.0}};
    for (long i = 0; i < n; i++) {
        double rhs = -rhs_in[i];
        if (stage == 0) {
            double u = ui[i];
            ui[i] = al[0][0] * u + al[0][1] * rhs * dt;
        } else {
            double u = un[i];
            double u1 = ui[i];
            ui[i] = al[stage][0] * u + al[stage][1] * (rhs * dt + u1);
        }
    }
}

/* ---------------------------------------------------------------------------
 * Pointwise physics
 * ------------------------------------------------------------------------- */
/* con2prim, src/physicsModel.cpp:331-344 */
void orc_con2prim(const double *q, double gamma, double mach, double *w)
{
    double rho = q[0];
    double u = q[1] / rho;
    double v = q[2] / rho;
    double ww = q[3] / rho;
    double E = q[4];
    double p = fabs((E - 0.5 * rho * (u * u + v * v + ww * ww)) * (gamma - 1.0));
    double e = p / rho / (gamma - 1.0);
    double alpha = sqrt(gamma * p / rho);
    double T = gamma * mach * mach * p / rho;
    w[0] = rho; w[1] = u; w[2] = v; w[3] = ww; w[4] = p; w[5] = T; w[6] = e; w[7] = alpha;
}

/* AUSM Mach/pressure splittings, src/physicsModel.cpp:1181-1228 */
static double machL(double mach)
{
    if (fabs(mach) < 1.0) return 0.25 * pow((mach + 1.0), 2) + 0.125 * pow((mach * mach - 1), 2);
    return 0.50 * (mach + fabs(mach));
}
static double machR(double mach)
{
    if (fabs(mach) < 1.0) return -0.25 * pow((mach - 1.0), 2) - 0.125 * pow((mach * mach - 1), 2);
    return 0.50 * (mach - fabs(mach));
}
static double presL(double mach, double p)
{
    if (fabs(mach) < 1.0) return p * (mach + 1.0) / 2.0;
    return 0.50 * p * (mach + fabs(mach)) / mach;
}
static double presR(double mach, double p)
{
    if (fabs(mach) < 1.0) return -p * (mach - 1.0) / 2.0;
    return 0.5 * p * (mach - fabs(mach)) / mach;
}

/* AUSM+ pressure splittings, src/physicsModel.cpp:1257-1277 (getMachLp/Rp at
 * :1231-1255 equal getMachL/R) */
static double presLp(double mach, double p)
{
    if (fabs(mach) < 1.0)
        return p * (0.25 * pow((mach + 1.0), 2.0) * (2.0 - mach) + 0.1875 * mach * pow((mach * mach - 1.0), 2.0));
    return 0.50 * p * (mach + fabs(mach)) / mach;
}
static double presRp(double mach, double p)
{
    if (fabs(mach) < 1.0)
        return p * (0.25 * pow(mach - 1.0, 2.0) * (2.0 + mach) - 0.1875 * mach * pow((mach * mach - 1.0), 2.0));
    return 0.5 * p * (mach - fabs(mach)) / mach;
}

/* Flux-vector splitting at one point in direction dir (0,1,2):
 * LF: src/physicsModel.cpp:540-565 (x), :587-619 (y), :641-674 (z);
 * AUSM: src/physicsModel.cpp:790-820 (x), :843-868 (y), :896-921 (z);
 * AUSM+: the same fluxes with the AUSM+ pressure splitting,
 * src/physicsModel.cpp:1040-1063 (x), :1093-1118 (y), :1148-1173 (z). */
void orc_split(int riemann, int dir, const double *q, const double *w, double *fp, double *fm)
{
    double ud = w[1 + dir], p0 = w[4], alpha0 = w[7];
    if (riemann == ORC_LF) {
        double lam = fabs(ud) + alpha0;
        double flux[5];
        flux[0] = ud * q[0];
        flux[1] = ud * q[1];
        flux[2] = ud * q[2];
        flux[3] = ud * q[3];
        flux[1 + dir] = ud * q[1 + dir] + p0;
        flux[4] = ud * (q[4] + p0);
        for (int v = 0; v < 5; v++) {
            fp[v] = 0.5 * (flux[v] + lam * q[v]);
            fm[v] = 0.5 * (flux[v] - lam * q[v]);
        }
    } else {
        double mach = ud / alpha0;
        double machp = machL(mach), machn = machR(mach);
        double p_p = (riemann == ORC_AUSMP) ? presLp(mach, p0) : presL(mach, p0);
        double p_n = (riemann == ORC_AUSMP) ? presRp(mach, p0) : presR(mach, p0);
        for (int v = 0; v < 4; v++) {
            fp[v] = machp * alpha0 * q[v];
            fm[v] = machn * alpha0 * q[v];
        }
        fp[1 + dir] = machp * alpha0 * q[1 + dir] + p_p;
        fm[1 + dir] = machn * alpha0 * q[1 + dir] + p_n;
        fp[4] = machp * alpha0 * q[4] + machp * alpha0 * p0;
        fm[4] = machn * alpha0 * q[4] + machn * alpha0 * p0;
    }
}

/* sutherland / constant, src/physicsModel.cpp:193-266 */
void orc_viscosity(int law, double T, double reynolds, double mach, double pt, double gamma,
                   double t_ref, double t_inf, double *miu, double *coeff)
{
    double non_dim_T_ref = t_ref / t_inf;
    double miu0;
    if (law == ORC_SUTHERLAND)
        miu0 = (1.0 + non_dim_T_ref) / (T + non_dim_T_ref) * pow(T, 1.5) / reynolds;
    else
        miu0 = 1.0 / reynolds;
    *miu = miu0;
    *coeff = miu0 / ((gamma - 1.0) * mach * mach * pt);
}

/* ---------------------------------------------------------------------------
 * Whole solver
 * ------------------------------------------------------------------------- */
struct orc_solver {
    orc_config c;
    long N;
    int n[3], np[3], nloc[3];
    double h[3];
    int ts[3];      /* inviscid halo per direction = tilesize */
    int tsv[3];     /* viscous tilesize: cd8 widened to 4 */
    int bf;         /* ghost width = max halos */
    double *q, *qn, *rhs;
    double *prim;   /* 8*N: rho,u,v,w,p,T,e,c */
    double *fv;     /* 15*N viscous fluxes (dir*5+v) */
    double *dinv;   /* 15*N inviscid derivatives */
    double *dvis;   /* 15*N viscous derivatives */
    double *grad;   /* 12*N: d(u,v,w,T)/d(x,y,z), index (f*3+dir) */
    double *lam;    /* per-rank max eigen (3*P) */
    double dt_global, dt_local, res, time;
};

static inline long gidx(const orc_solver *s, int i, int j, int k)
{
    return (long)i + (long)s->n[0] * ((long)j + (long)s->n[1] * k);
}

/* Run a per-line operation along direction d over every global line; for each
 * rank segment the callback sees the extended periodic line. */
typedef void (*line_fn)(orc_solver *s, int d, const double *ext, int seg, double *res_line, void *ctx);

static void line_coords(const orc_solver *s, int d, long line, int *a, int *b)
{
    int d1 = (d + 1) % 3, d2 = (d + 2) % 3;
    (void)d1; (void)d2;
    if (d == 0) { *a = (int)(line % s->n[1]); *b = (int)(line / s->n[1]); }      /* j, k */
    else if (d == 1) { *a = (int)(line % s->n[2]); *b = (int)(line / s->n[2]); } /* k, i */
    else { *a = (int)(line % s->n[0]); *b = (int)(line / s->n[0]); }             /* i, j */
}

static inline long line_idx(const orc_solver *s, int d, int a, int b, int t)
{
    if (d == 0) return gidx(s, t, a, b);
    if (d == 1) return gidx(s, b, t, a);
    return gidx(s, a, b, t);
}

/* derivative of a field along direction d with the viscous scheme (rank-local
 * ghosts from periodic neighbours), written to out (global layout) */
static void diff_dir(orc_solver *s, int d, const double *in, double *out)
{
    int n = s->n[d], nl = s->nloc[d], bf = s->bf;
    long nlines = s->N / n;
#pragma omp parallel num_threads(s->c.nthreads)
    {
        double *ext = (double *)malloc(sizeof(double) * (n + 2 * bf + 2 * nl));
        double *res = (double *)malloc(sizeof(double) * nl);
#pragma omp for schedule(static)
        for (long L = 0; L < nlines; L++) {
            int a, b;
            line_coords(s, d, L, &a, &b);
            for (int t = -bf; t < n + bf; t++) {
                int tt = ((t % n) + n) % n;
                ext[t + bf] = in[line_idx(s, d, a, b, tt)];
            }
            for (int r = 0; r < s->np[d]; r++) {
                orc_line_diff(s->c.vis_scheme[d], ext + bf + r * nl, nl, s->tsv[d], s->h[d], res);
                for (int t = 0; t < nl; t++) out[line_idx(s, d, a, b, r * nl + t)] = res[t];
            }
        }
        free(ext);
        free(res);
    }
}

orc_solver *orc_create(const orc_config *cfg)
{
    orc_solver *s = (orc_solver *)calloc(1, sizeof(orc_solver));
    s->c = *cfg;
    if (s->c.nthreads < 1) s->c.nthreads = 1;
    s->n[0] = cfg->nx; s->n[1] = cfg->ny; s->n[2] = cfg->nz;
    s->np[0] = cfg->npx; s->np[1] = cfg->npy; s->np[2] = cfg->npz;
    double L[3] = {cfg->lx, cfg->ly, cfg->lz};
    s->N = (long)cfg->nx * cfg->ny * cfg->nz;
    for (int d = 0; d < 3; d++) {
        s->nloc[d] = s->n[d] / s->np[d];
        s->h[d] = L[d] / s->n[d];                 /* dx = xl/nx_global, src/block.cpp:62-64 */
        s->ts[d] = orc_scheme_halo(cfg->scheme[d]);
    }
    s->bf = 0;
    for (int d = 0; d < 3; d++) {
        /* The reference passes the inviscid halo as the viscous tilesize
         * (src/fieldOperator.cpp:520-526) and sizes the ghost strips by the
         * inviscid halos only (src/block.cpp:170-171), so cd8 behind a 3-wide
         * scheme, and cd6 behind weno3js or upwind1st, read outside their
         * strips.  Widened here to the viscous stencil's reach. */
        int need = vis_reach(cfg->vis_scheme[d]);
        s->tsv[d] = s->ts[d] < need ? need : s->ts[d];
        if (s->tsv[d] > s->bf) s->bf = s->tsv[d];
    }
    size_t N = (size_t)s->N;
    s->q = (double *)calloc(N * 5, sizeof(double));
    s->qn = (double *)calloc(N * 5, sizeof(double));
    s->rhs = (double *)calloc(N * 5, sizeof(double));
    s->prim = (double *)calloc(N * 8, sizeof(double));
    s->fv = (double *)calloc(N * 15, sizeof(double));
    s->dinv = (double *)calloc(N * 15, sizeof(double));
    s->dvis = (double *)calloc(N * 15, sizeof(double));
    s->grad = (double *)calloc(N * 12, sizeof(double));
    s->lam = (double *)calloc(3 * (size_t)cfg->npx * cfg->npy * cfg->npz, sizeof(double));
    return s;
}

void orc_destroy(orc_solver *s)
{
    if (!s) return;
    free(s->q); free(s->qn); free(s->rhs); free(s->prim); free(s->fv);
    free(s->dinv); free(s->dvis); free(s->grad); free(s->lam);
    free(s);
}

void orc_set_state(orc_solver *s, const double *q) { memcpy(s->q, q, sizeof(double) * 5 * s->N); }
void orc_get_state(const orc_solver *s, double *q) { memcpy(q, s->q, sizeof(double) * 5 * s->N); }

static int rank_of(const orc_solver *s, int i, int j, int k)
{
    int px = i / s->nloc[0], py = j / s->nloc[1], pz = k / s->nloc[2];
    return px + s->np[0] * (py + s->np[1] * pz);
}

/* inviscid sweep along d for all 5 variables: split fluxes, reconstruct both
 * signs per rank segment, combine and difference (src/euler3d.cpp:23-56,
 * 132-253, 307-387) */
static void inviscid_dir(orc_solver *s, int d)
{
    int n = s->n[d], nl = s->nloc[d], bf = s->bf;
    long nlines = s->N / n;
    const double gamma = s->c.gamma;
    (void)gamma;
#pragma omp parallel num_threads(s->c.nthreads)
    {
        double *fpx = (double *)malloc(sizeof(double) * 5 * (n + 2 * bf));
        double *fmx = (double *)malloc(sizeof(double) * 5 * (n + 2 * bf));
        double *rp = (double *)malloc(sizeof(double) * (nl + 1));
        double *rm = (double *)malloc(sizeof(double) * (nl + 1));
#pragma omp for schedule(static)
        for (long L = 0; L < nlines; L++) {
            int a, b;
            line_coords(s, d, L, &a, &b);
            for (int t = -bf; t < n + bf; t++) {
                int tt = ((t % n) + n) % n;
                long id = line_idx(s, d, a, b, tt);
                double qq[5], w[8], fp[5], fm[5];
                for (int v = 0; v < 5; v++) qq[v] = s->q[v * s->N + id];
                for (int v = 0; v < 8; v++) w[v] = s->prim[v * s->N + id];
                orc_split(s->c.riemann, d, qq, w, fp, fm);
                for (int v = 0; v < 5; v++) {
                    fpx[v * (n + 2 * bf) + t + bf] = fp[v];
                    fmx[v * (n + 2 * bf) + t + bf] = fm[v];
                }
            }
            for (int v = 0; v < 5; v++) {
                for (int r = 0; r < s->np[d]; r++) {
                    const double *ep = fpx + v * (n + 2 * bf) + bf + r * nl;
                    const double *em = fmx + v * (n + 2 * bf) + bf + r * nl;
                    orc_line_reconstruct(s->c.scheme[d], 1, ep, nl, s->ts[d], rp);
                    orc_line_reconstruct(s->c.scheme[d], -1, em, nl, s->ts[d], rm);
                    for (int t = 0; t < nl; t++) {
                        double f1 = rp[t + 1] + rm[t + 1];
                        double f0 = rp[t] + rm[t];
                        double df_loc = f1 - f0;
                        s->dinv[(size_t)(d * 5 + v) * s->N + line_idx(s, d, a, b, r * nl + t)] = df_loc / s->h[d];
                    }
                }
            }
        }
        free(fpx); free(fmx); free(rp); free(rm);
    }
}

static void compute_prims_and_eigen(orc_solver *s)
{
    int P = s->np[0] * s->np[1] * s->np[2];
    for (int r = 0; r < 3 * P; r++) s->lam[r] = 0.0;
    for (int k = 0; k < s->n[2]; k++)
        for (int j = 0; j < s->n[1]; j++)
            for (int i = 0; i < s->n[0]; i++) {
                long id = gidx(s, i, j, k);
                double qq[5], w[8];
                for (int v = 0; v < 5; v++) qq[v] = s->q[v * s->N + id];
                orc_con2prim(qq, s->c.gamma, s->c.mach, w);
                for (int v = 0; v < 8; v++) s->prim[v * s->N + id] = w[v];
                int r = rank_of(s, i, j, k);
                /* max eigenvalue per direction; LF stores the z value into the
                 * eta slot (src/physicsModel.cpp:653-654) */
                double lx = fabs(w[1]) + w[7], ly = fabs(w[2]) + w[7], lz = fabs(w[3]) + w[7];
                s->lam[3 * r + 0] = fmax(lx, s->lam[3 * r + 0]);
                s->lam[3 * r + 1] = fmax(ly, s->lam[3 * r + 1]);
                if (s->c.riemann == ORC_LF) s->lam[3 * r + 1] = fmax(lz, s->lam[3 * r + 1]);
                else s->lam[3 * r + 2] = fmax(lz, s->lam[3 * r + 2]);
            }
}

/* AddForcing_Const (src/physicsModel.cpp:1368-1677) evaluated with a direct
 * DFT restricted to the modes it can touch: the band |kmag-k_f|<0.5 plus k=0,
 * which enters the band energy sums through the kmag<1e-6 -> 1 substitution
 * (:1483-1484,1534).  Forward transform normalised by 1/N, backward not
 * (src/fft.cpp:280-281). */
typedef struct { int kx, ky, kz; } kmode;

static int band_modes(const orc_solver *s, kmode *out, int maxm)
{
    int cnt = 0, K = (int)ceil(s->c.forcing_k + 0.5) + 1;
    for (int kz = -K; kz <= K; kz++)
        for (int ky = -K; ky <= K; ky++)
            for (int kx = -K; kx <= K; kx++) {
                /* wave numbers as the reference folds them: index i >= n/2 -> i-n */
                int ix = ((kx % s->n[0]) + s->n[0]) % s->n[0];
                int iy = ((ky % s->n[1]) + s->n[1]) % s->n[1];
                int iz = ((kz % s->n[2]) + s->n[2]) % s->n[2];
                double kx0 = (ix >= s->n[0] / 2.0) ? ix - s->n[0] : ix;
                double ky0 = (iy >= s->n[1] / 2.0) ? iy - s->n[1] : iy;
                double kz0 = (iz >= s->n[2] / 2.0) ? iz - s->n[2] : iz;
                if (kx0 != kx || ky0 != ky || kz0 != kz) continue;
                double kmag = sqrt(kx0 * kx0 + ky0 * ky0 + kz0 * kz0);
                if (kmag < 1e-6) kmag = 1.0;
                if (fabs(kmag - s->c.forcing_k) < 0.5 && cnt < maxm) {
                    out[cnt].kx = kx; out[cnt].ky = ky; out[cnt].kz = kz; cnt++;
                }
            }
    return cnt;
}

static void add_forcing_const(orc_solver *s)
{
    kmode modes[512];
    int nm = band_modes(s, modes, 512);
    const int nx = s->n[0], ny = s->n[1], nz = s->n[2];
    const double twopi = 8.0 * atan(1.0);
    double *uh = (double *)calloc((size_t)nm * 6, sizeof(double)); /* (re,im) x 3 comps */
    /* forward DFT at band modes, separable sums in long double-free FP64 */
    for (int m = 0; m < nm; m++) {
        for (int c = 0; c < 3; c++) {
            const double *u = s->prim + (size_t)(1 + c) * s->N;
            double sre = 0.0, sim = 0.0;
            for (int k = 0; k < nz; k++) {
                double pre_k = 0.0, pim_k = 0.0;
                for (int j = 0; j < ny; j++) {
                    double pre_j = 0.0, pim_j = 0.0;
                    for (int i = 0; i < nx; i++) {
                        long ph = ((long)modes[m].kx * i) % nx; if (ph < 0) ph += nx;
                        double ang = -twopi * (double)ph / nx;
                        double val = u[gidx(s, i, j, k)];
                        pre_j += val * cos(ang);
                        pim_j += val * sin(ang);
                    }
                    long ph = ((long)modes[m].ky * j) % ny; if (ph < 0) ph += ny;
                    double ang = -twopi * (double)ph / ny, cs = cos(ang), sn = sin(ang);
                    pre_k += pre_j * cs - pim_j * sn;
                    pim_k += pre_j * sn + pim_j * cs;
                }
                long ph = ((long)modes[m].kz * k) % nz; if (ph < 0) ph += nz;
                double ang = -twopi * (double)ph / nz, cs = cos(ang), sn = sin(ang);
                sre += pre_k * cs - pim_k * sn;
                sim += pre_k * sn + pim_k * cs;
            }
            double Ntot = (double)nx * ny * nz;
            uh[m * 6 + 2 * c] = sre / Ntot;
            uh[m * 6 + 2 * c + 1] = sim / Ntot;
        }
    }
    /* projection, helical mix, band energies (src/physicsModel.cpp:1476-1557) */
    double sigma = s->c.forcing_amp_hel, Ek_s_k = 0.0, Ek_c_k = 0.0;
    double *vs = (double *)calloc((size_t)nm * 6, sizeof(double));
    for (int m = 0; m < nm; m++) {
        double kx0 = modes[m].kx, ky0 = modes[m].ky, kz0 = modes[m].kz;
        double kmag = sqrt(kx0 * kx0 + ky0 * ky0 + kz0 * kz0);
        if (kmag < 1e-6) kmag = 1.0;
        double ure = uh[m * 6 + 0], uim = uh[m * 6 + 1], vre = uh[m * 6 + 2], vim = uh[m * 6 + 3];
        double wre = uh[m * 6 + 4], wim = uh[m * 6 + 5];
        double ku_re = (kx0 * ure + ky0 * vre + kz0 * wre) / kmag;
        double ku_im = (kx0 * uim + ky0 * vim + kz0 * wim) / kmag;
        double s_re[3], s_im[3], c_re[3], c_im[3];
        s_re[0] = ure - ku_re * kx0 / kmag; s_re[1] = vre - ku_re * ky0 / kmag; s_re[2] = wre - ku_re * kz0 / kmag;
        s_im[0] = uim - ku_im * kx0 / kmag; s_im[1] = vim - ku_im * ky0 / kmag; s_im[2] = wim - ku_im * kz0 / kmag;
        c_re[0] = ure - s_re[0]; c_re[1] = vre - s_re[1]; c_re[2] = wre - s_re[2];
        c_im[0] = uim - s_im[0]; c_im[1] = vim - s_im[1]; c_im[2] = wim - s_im[2];
        double kxu0_re = -(ky0 * s_im[2] - kz0 * s_im[1]) / kmag;
        double kxu1_re = (kx0 * s_im[2] - kz0 * s_im[0]) / kmag;
        double kxu2_re = -(kx0 * s_im[1] - ky0 * s_im[0]) / kmag;
        double kxu0_im = (ky0 * s_re[2] - kz0 * s_re[1]) / kmag;
        double kxu1_im = -(kx0 * s_re[2] - kz0 * s_re[0]) / kmag;
        double kxu2_im = (kx0 * s_re[1] - ky0 * s_re[0]) / kmag;
        s_re[0] = (s_re[0] + sigma * kxu0_re); s_re[1] = (s_re[1] + sigma * kxu1_re); s_re[2] = (s_re[2] + sigma * kxu2_re);
        s_im[0] = (s_im[0] + sigma * kxu0_im); s_im[1] = (s_im[1] + sigma * kxu1_im); s_im[2] = (s_im[2] + sigma * kxu2_im);
        Ek_s_k += 0.5 * (s_re[0] * s_re[0] + s_re[1] * s_re[1] + s_re[2] * s_re[2]
                         + s_im[0] * s_im[0] + s_im[1] * s_im[1] + s_im[2] * s_im[2]);
        Ek_c_k += 0.5 * (c_re[0] * c_re[0] + c_re[1] * c_re[1] + c_re[2] * c_re[2]
                         + c_im[0] * c_im[0] + c_im[1] * c_im[1] + c_im[2] * c_im[2]);
        for (int c = 0; c < 3; c++) { vs[m * 6 + 2 * c] = s_re[c]; vs[m * 6 + 2 * c + 1] = s_im[c]; }
    }
    double theta = sqrt(fabs(s->c.forcing_amp - Ek_c_k) / Ek_s_k);
    /* forcing spectrum (only |kmag-k_f|<0.5 without the k=0 substitution) and
     * unnormalised backward transform, real part (src/physicsModel.cpp:1560-1603) */
    double *force = (double *)calloc((size_t)s->N * 3, sizeof(double));
    for (int m = 0; m < nm; m++) {
        double kx0 = modes[m].kx, ky0 = modes[m].ky, kz0 = modes[m].kz;
        double kmag = sqrt(kx0 * kx0 + ky0 * ky0 + kz0 * kz0);
        if (!(fabs(kmag - s->c.forcing_k) < 0.5)) continue;
        for (int c = 0; c < 3; c++) {
            double fre = vs[m * 6 + 2 * c] * (theta - 1.0), fim = vs[m * 6 + 2 * c + 1] * (theta - 1.0);
            for (int k = 0; k < nz; k++)
                for (int j = 0; j < ny; j++)
                    for (int i = 0; i < nx; i++) {
                        long ph = ((long)modes[m].kx * i * (long)ny * nz + (long)modes[m].ky * j * (long)nx * nz
                                   + (long)modes[m].kz * k * (long)nx * ny) % ((long)nx * ny * nz);
                        if (ph < 0) ph += (long)nx * ny * nz;
                        double ang = twopi * (double)ph / ((double)nx * ny * nz);
                        force[(size_t)c * s->N + gidx(s, i, j, k)] += fre * cos(ang) - fim * sin(ang);
                    }
        }
    }
    /* energy-input correction and RHS update (src/physicsModel.cpp:1605-1667) */
    double A_sum = 0.0;
    for (long id = 0; id < s->N; id++) {
        double u0 = s->prim[1 * s->N + id], v0 = s->prim[2 * s->N + id], w0 = s->prim[3 * s->N + id];
        double fx = 1.0 * force[id], fy = 1.0 * force[s->N + id], fz = 1.0 * force[2 * s->N + id];
        A_sum += fx * u0 + fy * v0 + fz * w0;
    }
    double A_aver = A_sum / (double)s->N;
    for (long id = 0; id < s->N; id++) {
        double u0 = s->prim[1 * s->N + id], v0 = s->prim[2 * s->N + id], w0 = s->prim[3 * s->N + id];
        double fx = 1.0 * force[id], fy = 1.0 * force[s->N + id], fz = 1.0 * force[2 * s->N + id];
        s->rhs[1 * s->N + id] -= fx;
        s->rhs[2 * s->N + id] -= fy;
        s->rhs[3 * s->N + id] -= fz;
        s->rhs[4 * s->N + id] -= fx * u0 + fy * v0 + fz * w0 - A_aver;
    }
    free(uh); free(vs); free(force);
}

/* The analytic forcings and the rotating-frame sources, in the order of
 * physicsModel::solveSource (src/physicsModel.cpp:1280-1319): AddForcing_ABC
 * (:1679-1739) or AddForcing_TG (:1741-1799), then AddCoriolisForce
 * (:1801-1840), then AddCentrifugalForce (:1842-1913), whose energy term is
 * subtracted from the y-momentum RHS as written (:1900).  Positions are the
 * per-rank cell centres (src/block.cpp:116-118); u, v, w the stage input's. */
static void add_sources(orc_solver *s)
{
    const orc_config *c = &s->c;
    if (c->forcing != ORC_FORCE_ABC && c->forcing != ORC_FORCE_TG && !c->coriolis && !c->centrifugal) return;
    const long N = s->N;
    const double A = c->forcing_amp, kf = c->forcing_k;
    const double Ro = c->rossby, xc = c->x_center, yc = c->y_center, r0 = c->well_rad;
    const double dx = c->lx / c->nx, dy = c->ly / c->ny, dz = c->lz / c->nz;
    double *r1 = s->rhs + N, *r2 = s->rhs + 2 * N, *r3 = s->rhs + 3 * N, *r4 = s->rhs + 4 * N;
    for (int k = 0; k < s->n[2]; k++)
        for (int j = 0; j < s->n[1]; j++)
            for (int i = 0; i < s->n[0]; i++) {
                const long id = gidx(s, i, j, k);
                /* cell centres per rank minus the origin, src/block.cpp:116-118 */
                const double x0 = c->lx / c->npx * (i / s->nloc[0]) + dx * (0.5 + i % s->nloc[0]) - c->origin[0];
                const double y0 = c->ly / c->npy * (j / s->nloc[1]) + dy * (0.5 + j % s->nloc[1]) - c->origin[1];
                const double z0 = c->lz / c->npz * (k / s->nloc[2]) + dz * (0.5 + k % s->nloc[2]) - c->origin[2];
                const double rho = s->prim[id], u0 = s->prim[N + id], v0 = s->prim[2 * N + id], w0 = s->prim[3 * N + id];
                if (c->forcing == ORC_FORCE_ABC) {
                    const double rho0 = 1.0;
                    double fx = A * rho0 * (cos(kf * y0) + 1.1 * sin(kf * z0));
                    double fy = A * rho0 * (1.1 * cos(kf * z0) + 0.9 * sin(kf * x0));
                    double fz = A * rho0 * (0.9 * cos(kf * x0) + sin(kf * y0));
                    r1[id] -= fx; r2[id] -= fy; r3[id] -= fz;
                    r4[id] -= fx * u0 + fy * v0 + fz * w0;
                } else if (c->forcing == ORC_FORCE_TG) {
                    double fx = A * 2.0 * (sin(kf * x0) * cos(kf * y0) * sin(kf * z0));
                    double fy = -A * 2.0 * (cos(kf * x0) * sin(kf * y0) * sin(kf * z0));
                    r1[id] -= fx; r2[id] -= fy;
                    r4[id] -= (fx * u0 + fy * v0);
                }
                if (c->coriolis) {
                    r1[id] -= rho * Ro * v0;
                    r2[id] -= -rho * Ro * u0;
                }
                if (c->centrifugal) {
                    double r = sqrt((x0 - xc) * (x0 - xc) + (y0 - yc) * (y0 - yc));
                    double fx = rho * Ro * Ro * (x0 - xc) / (1.0 + pow(r / r0, 20)) / 4.0;
                    double fy = rho * Ro * Ro * (y0 - yc) / (1.0 + pow(r / r0, 20)) / 4.0;
                    r1[id] -= fx;
                    r2[id] -= fy;
                    r2[id] -= fx * u0 + fy * v0;
                }
            }
}

void orc_rhs(orc_solver *s, double *rhs_out, double *dt_local)
{
    const long N = s->N;
    compute_prims_and_eigen(s);
    for (int d = 0; d < 3; d++) inviscid_dir(s, d);
    int ns = (s->c.model == ORC_NS);
    if (ns) {
        /* gradients of u,v,w,T (src/NaiverStokes3d.cpp:87-175) */
        const double *fld[4] = {s->prim + 1 * N, s->prim + 2 * N, s->prim + 3 * N, s->prim + 5 * N};
        for (int f = 0; f < 4; f++)
            for (int d = 0; d < 3; d++) diff_dir(s, d, fld[f], s->grad + (size_t)(f * 3 + d) * N);
        /* viscous fluxes (src/NaiverStokes3d.cpp:231-302) */
        for (long id = 0; id < N; id++) {
            const double *g = s->grad;
            double ux = g[0 * N + id], uy = g[1 * N + id], uz = g[2 * N + id];
            double vx = g[3 * N + id], vy = g[4 * N + id], vz = g[5 * N + id];
            double wx = g[6 * N + id], wy = g[7 * N + id], wz = g[8 * N + id];
            double tx = g[9 * N + id], ty = g[10 * N + id], tz = g[11 * N + id];
            double miu0, coeff0;
            orc_viscosity(s->c.viscosity, s->prim[5 * N + id], s->c.reynolds, s->c.mach, s->c.prandtl,
                          s->c.gamma, s->c.t_ref, s->c.t_inf, &miu0, &coeff0);
            double grad = ux + vy + wz;
            double tau_xx = miu0 * (2.0 * ux - 2.0 / 3.0 * grad);
            double tau_xy = miu0 * (uy + vx - 2.0 / 3.0 * grad);
#ifdef ORC_PLANT_TAUXY
            tau_xy = miu0 * (uy + vx);
#endif
            double tau_xz = miu0 * (uz + wx - 2.0 / 3.0 * grad);
            double tau_yy = miu0 * (2.0 * vy - 2.0 / 3.0 * grad);
            double tau_yz = miu0 * (vz + wy - 2.0 / 3.0 * grad);
            double tau_zz = miu0 * (2.0 * wz - 2.0 / 3.0 * grad);
            double tau_tx = coeff0 * tx, tau_ty = coeff0 * ty, tau_tz = coeff0 * tz;
#ifdef ORC_PLANT_HEAT
            tau_ty = coeff0 * tz;
#endif
            double u = s->prim[1 * N + id], v = s->prim[2 * N + id], w = s->prim[3 * N + id];
            double *F = s->fv;
            F[0 * N + id] = 0.0; F[1 * N + id] = tau_xx; F[2 * N + id] = tau_xy; F[3 * N + id] = tau_xz;
            F[4 * N + id] = u * tau_xx + v * tau_xy + w * tau_xz + tau_tx;
            F[5 * N + id] = 0.0; F[6 * N + id] = tau_xy; F[7 * N + id] = tau_yy; F[8 * N + id] = tau_yz;
            F[9 * N + id] = u * tau_xy + v * tau_yy + w * tau_yz + tau_ty;
            F[10 * N + id] = 0.0; F[11 * N + id] = tau_xz; F[12 * N + id] = tau_yz; F[13 * N + id] = tau_zz;
            F[14 * N + id] = u * tau_xz + v * tau_yz + w * tau_zz + tau_tz;
#ifdef ORC_PLANT_EFLUX
            F[4 * N + id] = u * tau_xx + v * tau_xy + tau_tx;
#endif
        }
        for (int d = 0; d < 3; d++)
            for (int v = 0; v < 5; v++)
                diff_dir(s, d, s->fv + (size_t)(d * 5 + v) * N, s->dvis + (size_t)(d * 5 + v) * N);
    }
    /* RHS assembly in the reference's accumulation order
     * (src/NaiverStokes3d.cpp:372-500; src/euler3d.cpp:389-457) */
    for (int v = 0; v < 5; v++)
        for (long id = 0; id < N; id++) {
            double r = s->dinv[(size_t)(0 * 5 + v) * N + id];
            r += s->dinv[(size_t)(1 * 5 + v) * N + id];
            r += s->dinv[(size_t)(2 * 5 + v) * N + id];
            if (ns) {
                r -= s->dvis[(size_t)(0 * 5 + v) * N + id];
                r -= s->dvis[(size_t)(1 * 5 + v) * N + id];
                r -= s->dvis[(size_t)(2 * 5 + v) * N + id];
            }
            s->rhs[(size_t)v * N + id] = r;
        }
    if (s->c.forcing == ORC_FORCE_CONST) add_forcing_const(s);
    add_sources(s);
    /* getDt per rank, then allReduceMin (src/euler3d.cpp:459-466) */
    int P = s->np[0] * s->np[1] * s->np[2];
    double dtmin = 0.0;
    for (int r = 0; r < P; r++) {
        double dt_xi = s->h[0] / s->lam[3 * r + 0];
        double dt_eta = s->h[1] / s->lam[3 * r + 1];
        double dt_zeta = s->h[2] / s->lam[3 * r + 2];
        double dt = 1.0 / (1.0 / dt_xi + 1.0 / dt_eta + 1.0 / dt_zeta);
        if (r == 0 || dt < dtmin) dtmin = dt;
    }
    s->dt_local = dtmin;
    if (dt_local) *dt_local = dtmin;
    if (rhs_out) memcpy(rhs_out, s->rhs, sizeof(double) * 5 * N);
}

void orc_step(orc_solver *s, int nsteps, double *dt, double *res, double *time)
{
    const long N = s->N;
    for (int it = 0; it < nsteps; it++) {
        for (int stage = 0; stage < 3; stage++) {
            orc_rhs(s, NULL, NULL);
            /* PDEData3d::solve (src/PDEData3d.cpp:70-91) */
            if (stage == 0) {
                memcpy(s->qn, s->q, sizeof(double) * 5 * N);
                s->dt_global = s->dt_local * s->c.cfl;
            }
            orc_rk3(stage, s->rhs, s->qn, s->q, s->dt_global, 5 * N);
            if (stage == 2) {
                /* setRES: per-rank sums, then allReduceSum (src/PDEData3d.cpp:93-128) */
                int P = s->np[0] * s->np[1] * s->np[2];
                double total = 0.0;
                for (int r = 0; r < P; r++) {
                    int px = r % s->np[0], py = (r / s->np[0]) % s->np[1], pz = r / (s->np[0] * s->np[1]);
                    double part = 0.0;
                    for (int v = 0; v < 5; v++)
                        for (int k = pz * s->nloc[2]; k < (pz + 1) * s->nloc[2]; k++)
                            for (int j = py * s->nloc[1]; j < (py + 1) * s->nloc[1]; j++)
                                for (int i = px * s->nloc[0]; i < (px + 1) * s->nloc[0]; i++) {
                                    long id = gidx(s, i, j, k);
                                    part += pow((s->q[(size_t)v * N + id] - s->qn[(size_t)v * N + id]), 2);
                                }
                    total += part;
                }
                s->res = sqrt(total / ((double)N * 5)) / (s->dt_global);
            }
        }
        s->time += s->dt_global;
    }
    if (dt) *dt = s->dt_global;
    if (res) *res = s->res;
    if (time) *time = s->time;
}

void orc_init_tgv(const orc_config *c, double *q)
{
    int nl[3] = {c->nx / c->npx, c->ny / c->npy, c->nz / c->npz};
    double dx = c->lx / c->nx, dy = c->ly / c->ny, dz = c->lz / c->nz;
    long N = (long)c->nx * c->ny * c->nz;
    for (int k = 0; k < c->nz; k++)
        for (int j = 0; j < c->ny; j++)
            for (int i = 0; i < c->nx; i++) {
                /* cell centres per rank, src/block.cpp:116-118 */
                int cx = i / nl[0], cy = j / nl[1], cz = k / nl[2];
                int li = i % nl[0], lj = j % nl[1], lk = k % nl[2];
                double x0 = c->lx / c->npx * cx + dx * (0.5 + li) - c->origin[0];
                double y0 = c->ly / c->npy * cy + dy * (0.5 + lj) - c->origin[1];
                double z0 = c->lz / c->npz * cz + dz * (0.5 + lk) - c->origin[2];
                double gamma = c->gamma, mach = c->mach;
                /* src/block.cpp:377-391 */
                double rho0 = 1.0;
                double u0 = sin(x0) * cos(y0) * cos(z0);
                double v0 = -cos(x0) * sin(y0) * cos(z0);
                double w0 = 0.0;
                double p0 = 1.0 / (gamma * mach * mach) + ((cos(2.0 * z0) + 2.0) * (cos(2.0 * x0) + cos(2.0 * y0)) - 2.0) / 16.0;
                long id = (long)i + (long)c->nx * ((long)j + (long)c->ny * k);
                q[id] = rho0;
                q[N + id] = rho0 * u0;
                q[2 * N + id] = rho0 * v0;
                q[3 * N + id] = rho0 * w0;
                q[4 * N + id] = p0 / (gamma - 1.0) + 0.5 * rho0 * (u0 * u0 + v0 * v0 + w0 * w0);
            }
}

/* initial_default (src/block.cpp:285-314): fluid at rest.  u_init, v_init and
 * w_init are read with getValue, whose map does not hold them and which
 * PhysModel.in cannot set (an unknown word is an error), so they are 0. */
void orc_init_default(const orc_config *c, double *q)
{
    long N = (long)c->nx * c->ny * c->nz;
    double gamma = c->gamma, mach = c->mach;
    for (long id = 0; id < N; id++) {
        q[id] = 1.0;
        q[N + id] = 0.0;
        q[2 * N + id] = 0.0;
        q[3 * N + id] = 0.0;
        q[4 * N + id] = 1.0 / (mach * mach * gamma) / (gamma - 1.0);
    }
}

/* initial_ivc (src/block.cpp:316-357): the isentropic vortex centred at
 * (5, 5), strength b = 0.5, advected by u = 0.5; p = rho^gamma, unscaled by
 * the Mach number, and the density exponent 2.5 as written. */
void orc_init_ivc(const orc_config *c, double *q)
{
    int nl[3] = {c->nx / c->npx, c->ny / c->npy, c->nz / c->npz};
    double dx = c->lx / c->nx, dy = c->ly / c->ny;
    long N = (long)c->nx * c->ny * c->nz;
    double gamma = c->gamma;
    double pie = 4.0 * atan(1.0);
    double b = 0.5, x_c = 5.0, y_c = 5.0;
    for (int k = 0; k < c->nz; k++)
        for (int j = 0; j < c->ny; j++)
            for (int i = 0; i < c->nx; i++) {
                double x0 = c->lx / c->npx * (i / nl[0]) + dx * (0.5 + i % nl[0]) - c->origin[0];
                double y0 = c->ly / c->npy * (j / nl[1]) + dy * (0.5 + j % nl[1]) - c->origin[1];
                double r = sqrt(pow(x0 - x_c, 2) + pow(y0 - y_c, 2));
                double rho0 = pow((1 - (gamma - 1.0) * b * b * exp(1 - r * r) / (8.0 * gamma * pie * pie)), 2.5);
                double u0 = 0.5 - b / (2.0 * pie) * exp((1 - r * r) / 2) * (y0 - y_c);
                double v0 = b / (2.0 * pie) * exp((1 - r * r) / 2) * (x0 - x_c);
                double w0 = 0.0;
                double p0 = pow(rho0, gamma);
                long id = (long)i + (long)c->nx * ((long)j + (long)c->ny * k);
                q[id] = rho0;
                q[N + id] = rho0 * u0;
                q[2 * N + id] = rho0 * v0;
                q[3 * N + id] = rho0 * w0;
                q[4 * N + id] = p0 / (gamma - 1.0) + 0.5 * rho0 * (u0 * u0 + v0 * v0 + w0 * w0);
            }
}

void orc_stats(orc_solver *s, double *out)
{
    const long N = s->N;
    compute_prims_and_eigen(s);
    double *g = (double *)malloc(sizeof(double) * 9 * (size_t)N);
    for (int f = 0; f < 3; f++)
        for (int d = 0; d < 3; d++) diff_dir(s, d, s->prim + (size_t)(1 + f) * N, g + (size_t)(f * 3 + d) * N);
    double non_dim_T_ref = s->c.t_ref / s->c.t_inf;
    double m[16] = {0};
    double u_rms_aver = 0, miu_aver = 0, rho_bar = 0, div_aver = 0, alpha_aver = 0, Skew2 = 0, Skew3 = 0;
    /* src/postproc.cpp:796-842 */
    for (long id = 0; id < N; id++) {
        double ux = g[0 * N + id], uy = g[1 * N + id], uz = g[2 * N + id];
        double vx = g[3 * N + id], vy = g[4 * N + id], vz = g[5 * N + id];
        double wx = g[6 * N + id], wy = g[7 * N + id], wz = g[8 * N + id];
        double u0 = s->prim[1 * N + id], u1 = s->prim[2 * N + id], u2 = s->prim[3 * N + id];
        double rho = s->prim[id], T = s->prim[5 * N + id], alpha = s->prim[7 * N + id];
        double w0 = wy - vz, w1 = uz - wx, w2 = vx - uy;
        double q0 = ux * vy + ux * wz + vy * wz - uy * vx - uz * wx - vz * wy;
        double vals[16] = {ux, uy, uz, vx, vy, vz, wx, wy, wz, w0, w1, w2,
                           (w0 * w0 + w1 * w1 + w2 * w2) * 0.5, q0, w0 * u0 + w1 * u1 + w2 * u2,
                           0.5 * rho * (u0 * u0 + u1 * u1 + u2 * u2)};
        for (int t = 0; t < 16; t++) m[t] += vals[t];
        u_rms_aver += (u0 * u0 + u1 * u1 + u2 * u2) / 3.0;
        div_aver += (pow(ux, 2) + pow(vy, 2) + pow(wz, 2)) / 3.0;
        Skew3 += (pow(ux, 3) + pow(vy, 3) + pow(wz, 3)) / 3.0;
        Skew2 += (pow(ux, 2) + pow(vy, 2) + pow(wz, 2)) / 3.0;
        alpha_aver += alpha;
        miu_aver += (1.0 + non_dim_T_ref) / (T + non_dim_T_ref) * pow(T, 1.5);
        rho_bar += rho;
    }
    double Nd = (double)N;
    for (int t = 0; t < 16; t++) out[t] = m[t] / Nd;
    u_rms_aver /= Nd; div_aver /= Nd; miu_aver /= Nd; rho_bar /= Nd; Skew3 /= Nd; Skew2 /= Nd; alpha_aver /= Nd;
    double skew = Skew3 / pow(Skew2, 1.5);
#ifdef ORC_PLANT_SKEW
    skew = Skew3 / pow(Skew2, 2.0);
#endif
    double lambda = sqrt(u_rms_aver / div_aver);
#ifdef ORC_PLANT_LAMBDA
    lambda = u_rms_aver / div_aver;
#endif
    double Re_lambda = sqrt(u_rms_aver) * lambda * rho_bar / miu_aver * s->c.reynolds;
    double tau_t = lambda / sqrt(u_rms_aver);
    out[16] = sqrt(u_rms_aver);
    out[17] = sqrt(u_rms_aver * 3.0) / alpha_aver;
    out[18] = sqrt(div_aver);
    out[19] = miu_aver;
    out[20] = Re_lambda;
    out[21] = lambda;
    out[22] = rho_bar;
    out[23] = tau_t;
    out[24] = skew;
    free(g);
}

/* ---------------------------------------------------------------------------
 * Time-averaged statistics and TKE budget: postproc::solveAveraged0/1/2
 * (src/postproc.cpp:1186-2110) with the field lists of src/postproc.hpp:49-121.
 * type 1: primaries + gradients, 2: + derived, 3: + TKE fields and budget
 * (TimeAverageType, src/postproc.cpp:37-60).  Every derivative is solveGrad
 * (src/postproc.cpp:2111-2124,2284-2298): the viscous scheme per direction with
 * periodic rank ghosts, as diff_dir above.
 * ------------------------------------------------------------------------- */
struct orc_post {
    int type;
    long N;
    long step;                 /* averStep */
    double *tp, *ap;           /* primary: rho u v w p T e alpha miu coeff */
    double *tg, *ag;           /* grads (21) */
    double *td, *ad;           /* derived (28) */
    double *tt, *at;           /* TKE (49) */
    double *bud;               /* budget (45) */
};

static double *post_alloc(long n) { return (double *)calloc((size_t)n, sizeof(double)); }

orc_post *orc_post_create(const orc_solver *s, int type)
{
    if (type < 1 || type > 3) return NULL;
    orc_post *p = (orc_post *)calloc(1, sizeof(orc_post));
    long N = s->N;
    p->type = type;
    p->N = N;
    p->tp = post_alloc(ORC_NPRIM * N); p->ap = post_alloc(ORC_NPRIM * N);
    p->tg = post_alloc(ORC_NGRAD * N); p->ag = post_alloc(ORC_NGRAD * N);
    if (type >= 2) { p->td = post_alloc(ORC_NDER * N); p->ad = post_alloc(ORC_NDER * N); }
    if (type >= 3) { p->tt = post_alloc(ORC_NTKE * N); p->at = post_alloc(ORC_NTKE * N); p->bud = post_alloc(ORC_NBUD * N); }
    return p;
}

void orc_post_destroy(orc_post *p)
{
    if (!p) return;
    free(p->tp); free(p->ap); free(p->tg); free(p->ag); free(p->td); free(p->ad);
    free(p->tt); free(p->at); free(p->bud);
    free(p);
}

/* f_aver = (f_aver*averStep + f_temp)/(averStep + 1), src/postproc.cpp:1318 */
static void post_average(double *a, const double *t, long n, long step)
{
    for (long i = 0; i < n; i++) a[i] = (a[i] * step + t[i]) / (step + 1.0);
}

/* solveGrad of one field, one or all three directions */
static void post_grad(orc_solver *s, const double *f, double *dx, double *dy, double *dz)
{
    if (dx) diff_dir(s, 0, f, dx);
    if (dy) diff_dir(s, 1, f, dy);
    if (dz) diff_dir(s, 2, f, dz);
}

#define F(base, k) ((base) + (size_t)(k) * N)

/* solveAveraged0, src/postproc.cpp:1207-1349 */
static void post_level0(orc_post *p, orc_solver *s)
{
    const long N = s->N;
    compute_prims_and_eigen(s);
    double *tp = p->tp, *tg = p->tg;
    for (long id = 0; id < N; id++) {
        for (int v = 0; v < 8; v++) tp[v * N + id] = s->prim[v * N + id];   /* rho u v w p T e alpha */
        orc_viscosity(s->c.viscosity, s->prim[5 * N + id], s->c.reynolds, s->c.mach, s->c.prandtl, s->c.gamma,
                      s->c.t_ref, s->c.t_inf, &tp[8 * N + id], &tp[9 * N + id]);   /* getMiu: miu, coeff */
    }
    /* d(u, v, w, p, T)/d(x, y, z) */
    const int src[5] = {1, 2, 3, 4, 5};
    for (int f = 0; f < 5; f++) post_grad(s, F(tp, src[f]), F(tg, 3 * f), F(tg, 3 * f + 1), F(tg, 3 * f + 2));
    for (long id = 0; id < N; id++) {
        double uy = tg[1 * N + id], uz = tg[2 * N + id], vx = tg[3 * N + id], vz = tg[5 * N + id];
        double wx = tg[6 * N + id], wy = tg[7 * N + id];
        double o0 = wy - vz, o1 = uz - wx, o2 = vx - uy;
        tg[15 * N + id] = o0; tg[16 * N + id] = o1; tg[17 * N + id] = o2;
        tg[18 * N + id] = pow(o0, 2.0); tg[19 * N + id] = pow(o1, 2.0); tg[20 * N + id] = pow(o2, 2.0);
    }
    post_average(p->ap, tp, ORC_NPRIM * N, p->step);
    post_average(p->ag, tg, ORC_NGRAD * N, p->step);
}

/* solveAveraged1, src/postproc.cpp:1351-1483 */
static void post_level1(orc_post *p, orc_solver *s)
{
    const long N = s->N;
    const double *tp = p->tp, *tg = p->tg;
    double *td = p->td;
    for (long id = 0; id < N; id++) {
        double rho = tp[id], u = tp[N + id], v = tp[2 * N + id], w = tp[3 * N + id], e = tp[6 * N + id];
        double miu = tp[8 * N + id];
        double ux = tg[id], uy = tg[N + id], uz = tg[2 * N + id], vx = tg[3 * N + id], vy = tg[4 * N + id];
        double vz = tg[5 * N + id], wx = tg[6 * N + id], wy = tg[7 * N + id], wz = tg[8 * N + id];
        double o[28];
        o[0] = rho * u; o[1] = rho * v; o[2] = rho * w; o[3] = rho * e;
        o[4] = u * u; o[5] = u * v; o[6] = u * w; o[7] = v * v; o[8] = v * w; o[9] = w * w;
        o[10] = rho * u * u; o[11] = rho * u * v; o[12] = rho * u * w;
        o[13] = rho * v * v; o[14] = rho * v * w; o[15] = rho * w * w;
        double velgrad = ux + vy + wz;
        o[16] = miu * (2.0 * ux - 2.0 / 3.0 * velgrad);
        o[17] = miu * (uy + vx - 2.0 / 3.0 * velgrad);
        o[18] = miu * (uz + wx - 2.0 / 3.0 * velgrad);
        o[19] = miu * (2.0 * vy - 2.0 / 3.0 * velgrad);
        o[20] = miu * (vz + wy - 2.0 / 3.0 * velgrad);
        o[21] = miu * (2.0 * wz - 2.0 / 3.0 * velgrad);
        for (int t = 0; t < 6; t++) o[22 + t] = rho * o[16 + t];
        for (int t = 0; t < 28; t++) td[t * N + id] = o[t];
    }
    post_average(p->ad, td, ORC_NDER * N, p->step);
}

/* solveAveraged2, src/postproc.cpp:1485-2110 */
static void post_level2(orc_post *p, orc_solver *s)
{
    const long N = s->N;
    const double *tp = p->tp, *tg = p->tg, *td = p->td;
    double *tt = p->tt;
    /* derivatives of the viscous stress, src/postproc.cpp:1566-1592 */
    diff_dir(s, 0, F(td, 16), F(tt, 6));
    diff_dir(s, 0, F(td, 17), F(tt, 7));
    diff_dir(s, 0, F(td, 18), F(tt, 8));
    diff_dir(s, 1, F(td, 17), F(tt, 9));
    diff_dir(s, 1, F(td, 19), F(tt, 10));
    diff_dir(s, 1, F(td, 20), F(tt, 11));
    diff_dir(s, 2, F(td, 18), F(tt, 12));
    diff_dir(s, 2, F(td, 20), F(tt, 13));
    diff_dir(s, 2, F(td, 21), F(tt, 14));
    for (long id = 0; id < N; id++) {
        double rho = tp[id], u = tp[N + id], v = tp[2 * N + id], w = tp[3 * N + id], pr = tp[4 * N + id];
        double ux = tg[id], uy = tg[N + id], uz = tg[2 * N + id], vx = tg[3 * N + id], vy = tg[4 * N + id];
        double vz = tg[5 * N + id], wx = tg[6 * N + id], wy = tg[7 * N + id], wz = tg[8 * N + id];
        double txx = td[16 * N + id], txy = td[17 * N + id], txz = td[18 * N + id];
        double tyy = td[19 * N + id], tyz = td[20 * N + id], tzz = td[21 * N + id];
        tt[0 * N + id] = u * pr; tt[1 * N + id] = v * pr; tt[2 * N + id] = w * pr;
        tt[3 * N + id] = pr * ux; tt[4 * N + id] = pr * vy; tt[5 * N + id] = pr * wz;
        double a, b, c;
        a = u * txx; b = v * txy; c = w * txz;
        tt[15 * N + id] = a; tt[16 * N + id] = b; tt[17 * N + id] = c; tt[18 * N + id] = a + b + c;
        a = u * txy; b = v * tyy; c = w * tyz;
        tt[20 * N + id] = a; tt[21 * N + id] = b; tt[22 * N + id] = c; tt[23 * N + id] = a + b + c;
        a = u * txz; b = v * tyz; c = w * tzz;
        tt[25 * N + id] = a; tt[26 * N + id] = b; tt[27 * N + id] = c; tt[28 * N + id] = a + b + c;
        tt[30 * N + id] = txx * ux; tt[31 * N + id] = txy * vx; tt[32 * N + id] = txz * wx;
        tt[33 * N + id] = txy * uy; tt[34 * N + id] = tyy * vy; tt[35 * N + id] = tyz * wy;
        tt[36 * N + id] = txz * uz; tt[37 * N + id] = tyz * vz; tt[38 * N + id] = tzz * wz;
        tt[39 * N + id] = rho * u * u * u; tt[40 * N + id] = rho * u * u * v; tt[41 * N + id] = rho * u * u * w;
        tt[42 * N + id] = rho * u * v * v; tt[43 * N + id] = rho * u * v * w; tt[44 * N + id] = rho * u * w * w;
        tt[45 * N + id] = rho * v * v * v; tt[46 * N + id] = rho * v * v * w; tt[47 * N + id] = rho * v * w * w;
        tt[48 * N + id] = rho * w * w * w;
    }
    diff_dir(s, 0, F(tt, 18), F(tt, 19));
    diff_dir(s, 1, F(tt, 23), F(tt, 24));
    diff_dir(s, 2, F(tt, 28), F(tt, 29));
    post_average(p->at, tt, ORC_NTKE * N, p->step);

    /* the budget from the averaged fields, src/postproc.cpp:1720-2110 */
    const double *ap = p->ap, *ag = p->ag, *ad = p->ad, *at = p->at;
    double *bu = p->bud;
    post_grad(s, F(ad, 0), F(bu, 12), F(bu, 13), F(bu, 14));
    post_grad(s, F(ad, 1), F(bu, 15), F(bu, 16), F(bu, 17));
    post_grad(s, F(ad, 2), F(bu, 18), F(bu, 19), F(bu, 20));
    for (long id = 0; id < N; id++) {
        double pa = ap[4 * N + id];
        bu[21 * N + id] = at[0 * N + id] - ap[1 * N + id] * pa;
        bu[22 * N + id] = at[1 * N + id] - ap[2 * N + id] * pa;
        bu[23 * N + id] = at[2 * N + id] - ap[3 * N + id] * pa;
    }
    diff_dir(s, 0, F(bu, 21), F(bu, 24));
    diff_dir(s, 1, F(bu, 22), F(bu, 25));
    diff_dir(s, 2, F(bu, 23), F(bu, 26));
    for (long id = 0; id < N; id++) {
        double u = ap[1 * N + id], v = ap[2 * N + id], w = ap[3 * N + id];
        double txx = ad[16 * N + id], txy = ad[17 * N + id], txz = ad[18 * N + id];
        double tyy = ad[19 * N + id], tyz = ad[20 * N + id], tzz = ad[21 * N + id];
        bu[27 * N + id] = u * txx + v * txy + w * txz;
        bu[28 * N + id] = u * txy + v * tyy + w * tyz;
        bu[29 * N + id] = u * txz + v * tyz + w * tzz;
    }
    diff_dir(s, 0, F(bu, 27), F(bu, 30));
    diff_dir(s, 1, F(bu, 28), F(bu, 31));
    diff_dir(s, 2, F(bu, 29), F(bu, 32));
    for (long id = 0; id < N; id++) {
        double rho = ap[id], u = ap[1 * N + id], v = ap[2 * N + id], w = ap[3 * N + id], pa = ap[4 * N + id];
        double ux = ag[id], uy = ag[N + id], uz = ag[2 * N + id], vx = ag[3 * N + id], vy = ag[4 * N + id];
        double vz = ag[5 * N + id], wx = ag[6 * N + id], wy = ag[7 * N + id], wz = ag[8 * N + id];
        double px = ag[9 * N + id], py = ag[10 * N + id], pz = ag[11 * N + id];
        double ru = ad[0 * N + id], rv = ad[1 * N + id], rw = ad[2 * N + id];
        double ruu = ad[10 * N + id], ruv = ad[11 * N + id], ruw = ad[12 * N + id];
        double rvv = ad[13 * N + id], rvw = ad[14 * N + id], rww = ad[15 * N + id];
        double txx = ad[16 * N + id], txy = ad[17 * N + id], txz = ad[18 * N + id];
        double tyy = ad[19 * N + id], tyz = ad[20 * N + id], tzz = ad[21 * N + id];
        const double *d = bu;
        double tke = 0.5 * (ruu / rho - pow(ru / rho, 2.0)) + 0.5 * (rvv / rho - pow(rv / rho, 2.0)) +
                     0.5 * (rww / rho - pow(rw / rho, 2.0));
        double prod = (ruu - ru * ru / rho) * d[12 * N + id] + (ruv - ru * rv / rho) * d[13 * N + id] +
                      (ruw - ru * rw / rho) * d[14 * N + id] + (ruv - ru * rv / rho) * d[15 * N + id] +
                      (rvv - rv * rv / rho) * d[16 * N + id] + (rvw - rv * rw / rho) * d[17 * N + id] +
                      (ruw - ru * rw / rho) * d[18 * N + id] + (rvw - rv * rw / rho) * d[19 * N + id] +
                      (rww - rw * rw / rho) * d[20 * N + id];
        double p1 = -(u - ru / rho) * px - (v - rv / rho) * py - (w - rw / rho) * pz;
        double p2 = -d[24 * N + id] - d[25 * N + id] - d[26 * N + id];
        double p3 = (at[3 * N + id] - ux * pa) + (at[4 * N + id] - vy * pa) + (at[5 * N + id] - wz * pa);
        double v1 = (u - ru / rho) * at[6 * N + id] + (v - rv / rho) * at[7 * N + id] + (w - rw / rho) * at[8 * N + id] +
                    (u - ru / rho) * at[9 * N + id] + (v - rv / rho) * at[10 * N + id] + (w - rw / rho) * at[11 * N + id] +
                    (u - ru / rho) * at[12 * N + id] + (v - rv / rho) * at[13 * N + id] + (w - rw / rho) * at[14 * N + id];
        double v2 = (at[19 * N + id] - d[30 * N + id]) + (at[24 * N + id] - d[31 * N + id]) + (at[29 * N + id] - d[32 * N + id]);
        double v3 = -(at[30 * N + id] - ux * txx) - (at[31 * N + id] - vx * txy) - (at[32 * N + id] - wx * txz) -
                    (at[33 * N + id] - uy * txy) - (at[34 * N + id] - vy * tyy) - (at[35 * N + id] - wy * tyz) -
                    (at[36 * N + id] - uz * txz) - (at[37 * N + id] - vz * tyz) - (at[38 * N + id] - wz * tzz);
        bu[0 * N + id] = tke;
        bu[1 * N + id] = prod;
        bu[4 * N + id] = p1; bu[5 * N + id] = p2; bu[6 * N + id] = p3;
        bu[3 * N + id] = p1 + p2 + p3;
        bu[8 * N + id] = v1; bu[9 * N + id] = v2; bu[10 * N + id] = v3;
        bu[7 * N + id] = v1 + v2 + v3;
        bu[33 * N + id] = ru * tke; bu[34 * N + id] = rv * tke; bu[35 * N + id] = rw * tke;
    }
    diff_dir(s, 0, F(bu, 33), F(bu, 36));
    diff_dir(s, 1, F(bu, 34), F(bu, 37));
    diff_dir(s, 2, F(bu, 35), F(bu, 38));
    for (long id = 0; id < N; id++) {
        double rho = ap[id];
        double ru = ad[0 * N + id], rv = ad[1 * N + id], rw = ad[2 * N + id];
        double ruu = ad[10 * N + id], ruv = ad[11 * N + id], ruw = ad[12 * N + id];
        double rvv = ad[13 * N + id], rvw = ad[14 * N + id], rww = ad[15 * N + id];
        double r3[10];
        for (int t = 0; t < 10; t++) r3[t] = at[(39 + t) * N + id];   /* uuu uuv uuw uvv uvw uww vvv vvw vww www */
        bu[11 * N + id] = -bu[36 * N + id] - bu[37 * N + id] - bu[38 * N + id];
        /* the bracket of the last term, evaluated left to right as written */
        double kk = ruu - ru * ru / rho + rvv - rv * rv / rho + rww - rw * rw / rho;
        bu[39 * N + id] = 0.5 * r3[0] - ruu * ru / rho + 0.5 * ru * ru * ru / rho / rho +
                          0.5 * r3[3] - ruv * rv / rho + 0.5 * ru * rv * rv / rho / rho +
                          0.5 * r3[5] - ruw * rw / rho + 0.5 * ru * rw * rw / rho / rho - ru / rho * 0.5 * kk;
        bu[40 * N + id] = 0.5 * r3[1] - ruv * ru / rho + 0.5 * rv * ru * ru / rho / rho +
                          0.5 * r3[6] - rvv * rv / rho + 0.5 * rv * rv * rv / rho / rho +
                          0.5 * r3[8] - rvw * rw / rho + 0.5 * rv * rw * rw / rho / rho - rv / rho * 0.5 * kk;
        bu[41 * N + id] = 0.5 * r3[2] - ruw * ru / rho + 0.5 * rw * ru * ru / rho / rho +
                          0.5 * r3[7] - rvw * rv / rho + 0.5 * rw * rv * rv / rho / rho +
                          0.5 * r3[9] - rww * rw / rho + 0.5 * rw * rw * rw / rho / rho - rw / rho * 0.5 * kk;
    }
    diff_dir(s, 0, F(bu, 39), F(bu, 42));
    diff_dir(s, 1, F(bu, 40), F(bu, 43));
    diff_dir(s, 2, F(bu, 41), F(bu, 44));
    for (long id = 0; id < N; id++) bu[2 * N + id] = -bu[42 * N + id] - bu[43 * N + id] - bu[44 * N + id];
}
#undef F

void orc_post_update(orc_post *p, orc_solver *s)
{
    post_level0(p, s);
    if (p->type >= 2) post_level1(p, s);
    if (p->type >= 3) post_level2(p, s);
    p->step++;
}

const double *orc_post_field(const orc_post *p, int group, int index)
{
    static const int count[5] = {ORC_NPRIM, ORC_NGRAD, ORC_NDER, ORC_NTKE, ORC_NBUD};
    if (group < 0 || group > 4 || index < 0 || index >= count[group]) return NULL;
    const double *base[5] = {p->ap, p->ag, p->ad, p->at, p->bud};
    if (!base[group]) return NULL;
    return base[group] + (size_t)index * p->N;
}

long orc_post_steps(const orc_post *p) { return p->step; }
