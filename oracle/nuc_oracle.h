/*
 * nuc_oracle.h -- CPU restatement of the NUC3d per-stage right-hand side and
 * TVD-RK3 time advance (juncas/NUC3d-cartisian, src/).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it.  The product library never links or calls it.
 *
 * Data layout: every scalar field is a separate FP64 array, x fastest,
 * idx = i + nx*(j + ny*k) (src/field.cpp:96-100).  Conserved state is five
 * such arrays stored back to back: q[v*N + idx], v = rho, rhou, rhov, rhow, rhoE.
 *
 * Scheme ids follow the names accepted in inp/Method.in
 * (src/fieldOperator.cpp:128-385).
 */
#ifndef NUC_ORACLE_H
#define NUC_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_WENO5JS = 0, ORC_WENO5Z = 1, ORC_WENO7JS = 2, ORC_CRWENO5 = 3, ORC_HCCS = 4, ORC_WENO3JS = 5,
       ORC_UPWIND1ST = 6, ORC_HCCS78 = 7, ORC_HCCS77 = 8 };
enum { ORC_CD6 = 0, ORC_CD8 = 1, ORC_CD2 = 2 };
enum { ORC_LF = 0, ORC_AUSM = 1, ORC_AUSMP = 2 };
enum { ORC_EULER = 0, ORC_NS = 1 };
enum { ORC_SUTHERLAND = 0, ORC_CONSTANT = 1 };
enum { ORC_FORCE_NONE = 0, ORC_FORCE_CONST = 1, ORC_FORCE_ABC = 2, ORC_FORCE_TG = 3 };

typedef struct {
    int nx, ny, nz;          /* global grid */
    int npx, npy, npz;       /* emulated rank grid (inp/block.in line 1) */
    double lx, ly, lz;       /* box lengths (inp/block.in line 3) */
    int scheme[3];           /* inviscid reconstruction per direction */
    int vis_scheme[3];       /* viscous central difference per direction */
    int riemann, model, viscosity, forcing;
    double gamma, mach, reynolds, prandtl, t_ref, t_inf;
    double forcing_amp, forcing_amp_hel, forcing_k;
    double cfl;
    int nthreads;            /* OpenMP threads for line loops (CPU baseline) */
    int coriolis, centrifugal;                        /* source switches */
    double rossby, x_center, y_center, well_rad;      /* PhysModel.in */
    double origin[3];                                 /* block.in x/y/z_origin (src/block.cpp:61) */
} orc_config;

/* --- operator level: one line of n cells, f valid on [-bf, n+bf) --------- */
/* Halo width ("tilesize") of an inviscid scheme (src/fieldOperator.cpp:130-220). */
int orc_scheme_halo(int scheme);
/* Reconstruct interface values out[0..n] (out[m] = interface between cells m-1
 * and m) exactly as interpolationInner + interpolationBoundaryL/R do for one
 * line of a rank-local field of length n with ghost cells from its neighbours. */
void orc_line_reconstruct(int scheme, int upwind, const double *f, int n, int ts, double *out);
/* First derivative out[0..n-1] (differentialInner + differentialBoundaryL/R). */
void orc_line_diff(int scheme, const double *f, int n, int ts, double h, double *out);

/* --- reference-shaped Field operators (sweep along the first axis) -------- */
/* in: (n1,n2,n3) x-fastest; gL/gR: (bf,n2,n3) ghost fields; out: (n1+1,n2,n3). */
void orc_field_reconstruct(int scheme, int upwind, int ts, const double *in, int n1, int n2, int n3,
                           const double *gL, const double *gR, int bf, double *out);
void orc_field_diff(int scheme, int ts, const double *in, double h, int n1, int n2, int n3,
                    const double *gL, const double *gR, int bf, double *out);
void orc_rk3(int stage, const double *rhs, const double *un, double *ui, double dt, long n);

/* One interior row of hccs (src/hccs.cpp:138-183) from the 5-point upwind
 * stencil f[0..4]: out = {lo, di, hi, rhs} (sub-, main, super-diagonal, rhs). */
void orc_hccs_row(const double *f, double *out);

/* --- pointwise physics (src/physicsModel.cpp) ----------------------------- */
void orc_con2prim(const double *q, double gamma, double mach, double *prim /*rho,u,v,w,p,T,e,c*/);
void orc_split(int riemann, int dir, const double *q, const double *prim, double *fp, double *fm);
void orc_viscosity(int law, double T, double reynolds, double mach, double prandtl, double gamma,
                   double t_ref, double t_inf, double *miu, double *coeff);

/* --- whole solver --------------------------------------------------------- */
typedef struct orc_solver orc_solver;
orc_solver *orc_create(const orc_config *cfg);
void orc_destroy(orc_solver *s);
void orc_set_state(orc_solver *s, const double *q);
void orc_get_state(const orc_solver *s, double *q);
/* One right-hand-side evaluation of the current state (NaiverStokesData3d::solve /
 * EulerData3D::solve).  rhs: 5*N (may be NULL); dt_local: min over ranks. */
void orc_rhs(orc_solver *s, double *rhs, double *dt_local);
/* nsteps full RK3 steps (block::solve).  Returns the last step's dt and residual. */
void orc_step(orc_solver *s, int nsteps, double *dt, double *res, double *time);
/* Taylor-Green initial condition of src/block.cpp:359-395 on the rank grid. */
void orc_init_tgv(const orc_config *cfg, double *q);
void orc_init_default(const orc_config *c, double *q);
void orc_init_ivc(const orc_config *c, double *q);
/* Global statistics of postproc::solveQ (src/postproc.cpp:736-929):
 * out[0..15] = means of ux..wz(9), omega(3), enstrophy, Q, helicity, KE;
 * out[16..24] = u_rms, M_t, sqrt(div_aver), miu_aver, Re_lambda, lambda, rho_bar, tau_t, skew. */
void orc_stats(orc_solver *s, double *out);

/* --- time-averaged statistics and TKE budget ------------------------------
 * postproc::solveAveraged0/1/2 (src/postproc.cpp:1186-2110).  type 1: primaries
 * (10: rho u v w p T e alpha miu coeff) and gradients (21); 2: + derived (28);
 * 3: + TKE fields (49) and the budget (45); field order of src/postproc.hpp:49-121.
 * orc_post_update averages the solver's current state (averStep += 1).
 * orc_post_field: group 0 primary, 1 grads, 2 derived, 3 TKE (averaged), 4 budget. */
enum { ORC_NPRIM = 10, ORC_NGRAD = 21, ORC_NDER = 28, ORC_NTKE = 49, ORC_NBUD = 45 };
typedef struct orc_post orc_post;
orc_post *orc_post_create(const orc_solver *s, int type);
void orc_post_destroy(orc_post *p);
void orc_post_update(orc_post *p, orc_solver *s);
const double *orc_post_field(const orc_post *p, int group, int index);
long orc_post_steps(const orc_post *p);

#ifdef __cplusplus
}
#endif
#endif
