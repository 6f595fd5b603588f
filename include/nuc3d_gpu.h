/*
 * nuc3d_gpu.h -- C ABI of the MI355X (gfx950) implementation of NUC3d's
 * per-stage right-hand side and TVD-RK3 time advance.
 *
 * Plain C: pointers, sizes and status codes only.  Every call returns an int
 * status (NUC_OK on success) and never exits the process; nuc_last_error()
 * gives the message of the last failure on a context.
 *
 * Layout contract (src/field.cpp:96-100): each scalar field is an FP64 array
 * with x fastest, idx = i + nx*(j + ny*k).  A conserved state is five such
 * arrays back to back: rho, rho*u, rho*v, rho*w, rho*E.  Reconstructed
 * interface arrays hold n1+1 values per line, out[m] being the interface
 * between cells m-1 and m (src/weno5js.cpp:108).  Ghost strips are
 * (bf, n2, n3) fields: left ghost g <-> cell g-bf, right ghost g <-> cell n1+g
 * (src/bufferData.cpp:56-192).
 */
#ifndef NUC3D_GPU_H
#define NUC3D_GPU_H
#ifdef __cplusplus
extern "C" {
#endif

enum nuc_status {
    NUC_OK = 0,
    NUC_EINVAL = 1,       /* bad argument or unsupported configuration */
    NUC_ENOMEM = 2,       /* device allocation failed */
    NUC_EHIP = 3,         /* a HIP runtime call failed */
    NUC_ENONFINITE = 4,   /* state became NaN/Inf */
    NUC_ENODEV = 5,       /* no usable gfx950 device */
    NUC_EBUSY = 6         /* work still in flight (nuc_query, serving results) */
};

/* inviscid schemes, names as accepted in inp/Method.in (src/fieldOperator.cpp:128-220) */
enum nuc_scheme { NUC_WENO5JS = 0, NUC_WENO5Z = 1, NUC_WENO7JS = 2, NUC_CRWENO5 = 3, NUC_HCCS = 4,
                  NUC_WENO3JS = 5, NUC_UPWIND1ST = 6, NUC_HCCS78 = 7, NUC_HCCS77 = 8 };
/* viscous first-derivative schemes (src/fieldOperator.cpp:298-345) */
enum nuc_vis_scheme { NUC_CD6 = 0, NUC_CD8 = 1, NUC_CD2 = 2 };
/* flux-vector splitting (src/physicsModel.cpp:16-20) */
enum nuc_riemann { NUC_LF = 0, NUC_AUSM = 1, NUC_AUSMP = 2 };
/* PDE model (src/block.cpp:146-157) */
enum nuc_model { NUC_EULER3D = 0, NUC_NAVIERSTOKES3D = 1 };
/* viscosity law (src/physicsModel.cpp:45-49) */
enum nuc_viscosity { NUC_SUTHERLAND = 0, NUC_CONSTANT = 1 };
/* forcing (src/physicsModel.cpp:50-56): Const (FFT band forcing), ABC and TG
 * (analytic fields at the cell centres, src/physicsModel.cpp:1679-1799) */
enum nuc_forcing { NUC_FORCING_NONE = 0, NUC_FORCING_CONST = 1, NUC_FORCING_ABC = 2, NUC_FORCING_TG = 3 };
/* precision.  NUC_FP64: everything FP64.  NUC_MIXED: FP32 split fluxes,
 * stencils and viscous fluxes, FP64 conserved state, tridiagonal solves, RHS
 * accumulation and RK3 update (hccs and hccs77 keep FP64 stencils; a context
 * that uses only them runs FP64 throughout).  NUC_FP32: the mixed arithmetic with the
 * conserved state and the RHS also stored in FP32 (half their memory
 * traffic); uploads, downloads, statistics, spectra and averages see the
 * state widened to FP64.  NUC_FP32 does not take hccs or hccs77 (their
 * blend needs FP64 stencils), forcing or rotating-frame sources, packs or
 * nuc_rhs_noalloc (NUC_EINVAL). */
enum nuc_precision { NUC_FP64 = 0, NUC_MIXED = 1, NUC_FP32 = 2 };

typedef struct nuc_config {
    int nx, ny, nz;           /* global grid (inp/block.in line 2) */
    int npx, npy, npz;        /* rank grid the reference would run on (inp/block.in line 1);
                                 it places the per-rank closures of the schemes */
    double lx, ly, lz;        /* box lengths (inp/block.in line 3) */
    int scheme[3];            /* enum nuc_scheme per direction */
    int vis_scheme[3];        /* enum nuc_vis_scheme per direction */
    int riemann, model, viscosity, forcing, precision;
    double gamma, mach, reynolds, prandtl, t_ref, t_inf;
    double forcing_amp, forcing_amp_hel, forcing_k;
    double cfl;
    /* rotating-frame sources of PhysModel.in (src/physicsModel.cpp:1280-1319,
     * 1801-1913): Coriolis / Centrifugal switches (0 = "no"), Rossby number,
     * centre and radius of the centrifugal well */
    int coriolis, centrifugal;
    double rossby, x_center, y_center, well_rad;
    /* x/y/z_origin of inp/block.in line 4 (src/block.cpp:61): subtracted from
     * every cell centre (src/block.cpp:116-118), so it shifts the analytic
     * forcings, the centrifugal well and the analytic initial conditions */
    double origin[3];
} nuc_config;

typedef struct nuc_step_info {
    double dt;      /* dt_global of the last step (src/PDEData3d.cpp:76-80) */
    double time;    /* accumulated time (src/block.cpp:208-209) */
    double res;     /* residual of the last step (src/PDEData3d.cpp:93-128) */
    long step;      /* steps taken on this context */
} nuc_step_info;

typedef struct nuc_ctx nuc_ctx;

/* ---- context --------------------------------------------------------------
 * Replaces the per-rank state the reference's block owns (src/block.cpp:143-187):
 * PDEData3d (Q, Q_Euler, Q_work, RHS), EulerData3D/NaiverStokesData3d work
 * fields and the halo buffers.  All device memory is owned by the context.
 * Grids: each size divisible by its rank count (src/block.cpp:56-58), at most
 * 2^29 cells; lines above 512 cells need, for the explicit schemes, a divisor
 * of the length in [64, 255], for the compact ones (CRWENO5, hccs, hccs77)
 * rank blocks of at most 511 cells (else NUC_EINVAL). */
int nuc_create(const nuc_config *cfg, nuc_ctx **out);
int nuc_destroy(nuc_ctx *ctx);
/* the message of the last failure on ctx; with ctx NULL, the message of the
 * last failed nuc_create on the calling thread ("" when it succeeded) */
const char *nuc_last_error(const nuc_ctx *ctx);
/* bytes of device memory a context of this configuration allocates */
long long nuc_footprint(const nuc_config *cfg);

/* ---- state ---------------------------------------------------------------
 * q: 5*nx*ny*nz host doubles (the restart payload of src/block.cpp:640-671). */
int nuc_upload_state(nuc_ctx *ctx, const double *q);
int nuc_download_state(nuc_ctx *ctx, double *q);

/* ---- hot path ------------------------------------------------------------
 * nuc_rhs: one right-hand-side evaluation of the current state, i.e.
 * NaiverStokesData3d::solve / EulerData3D::solve (src/NaiverStokes3d.cpp:37-73,
 * src/euler3d.cpp:260-277) with the on-device periodic wrap in place of
 * MPIComunicator.  rhs_host (5N, may be NULL) receives RHS with dQ/dt = -RHS;
 * dt_local (may be NULL) the min over ranks of getDt (src/euler3d.cpp:459-466). */
int nuc_rhs(nuc_ctx *ctx, double *rhs_host, double *dt_local);
/* nuc_rhs_noalloc: nuc_rhs without a context and without any device
 * allocation.  No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under hipGraph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  q_dev (5N device doubles, SoA, x
 * fastest) is read, rhs_dev (5N device doubles) receives RHS with
 * dQ/dt = -RHS, dt_local_dev (one device double, may be NULL) the min over
 * ranks of getDt; all work is enqueued on `stream` (a hipStream_t, NULL for
 * the null stream) and the call returns without waiting.  The intermediates
 * (viscous fluxes, eigenvalue maxima) live in one fixed device array of the
 * library sized for NUC_NOALLOC_MAX_CELLS cells (72 MiB, held on every GPU
 * the library runs on until the process ends); larger grids and Const forcing
 * (whose transform plan owns buffers) are refused with NUC_EINVAL.  Calls that
 * overlap on one GPU share that array and race: serialise them (one stream,
 * or events).  Errors: nuc_last_error(NULL) on the calling thread. */
#define NUC_NOALLOC_MAX_CELLS 1048576L
int nuc_rhs_noalloc(const nuc_config *cfg, const double *q_dev, double *rhs_dev, double *dt_local_dev, void *stream);
/* nsteps TVD-RK3 steps: block::solve (src/block.cpp:191-215), i.e. three
 * (RHS + PDEData3d::solve) stages each.  info may be NULL. */
int nuc_step(nuc_ctx *ctx, int nsteps, nuc_step_info *info);
/* One RK stage: the RHS of the current stage state and PDEData3d::solve's
 * update (src/PDEData3d.cpp:70-91, src/TVD-RK3.cpp:16-157), i.e. one pass of
 * block::solve's stage loop (src/block.cpp:198-202).  Stages run in order
 * 0, 1, 2; dt is set at stage 0.  Between stages nuc_download_state,
 * nuc_stats and nuc_rhs see the current stage state; nuc_step and packs refuse
 * a context in the middle of a step, and nuc_upload_state starts a new step. */
int nuc_rk3_stage(nuc_ctx *ctx, int stage);
/* enqueue nsteps steps without waiting (timing); nuc_sync waits */
int nuc_step_async(nuc_ctx *ctx, int nsteps);
/* Replay steps from one captured hipGraph of a whole RK3 step (kernel nodes
 * only) instead of launching each kernel: removes the per-launch host cost that
 * dominates small boxes.  No counterpart in the reference (its step loop is
 * src/block.cpp:191-215).  Profiling (nuc_profile) falls back to launches. */
int nuc_set_graph(nuc_ctx *ctx, int enable);
/* Autotuner: time each candidate tile shape of the gradient kernel (planes per
 * workgroup of its z march: 4, 8, 16, 32) on ctx's current state and keep the
 * fastest; the choice is cached per (nx, ny, nz, precision) for the process.
 * Results do not depend on the tile shape.  No counterpart in the reference.
 * nuc_tile_shape reports the shape in use; nuc_set_tile_shape sets it
 * (1 <= planes <= nz). */
int nuc_autotune(nuc_ctx *ctx);
int nuc_tile_shape(const nuc_ctx *ctx, int *grad_planes);
int nuc_set_tile_shape(nuc_ctx *ctx, int grad_planes);
/* NUC_OK when all enqueued work of ctx has finished, NUC_EBUSY otherwise; never blocks. */
int nuc_query(nuc_ctx *ctx);
int nuc_sync(nuc_ctx *ctx, nuc_step_info *info);

/* ---- packs: several small boxes in one launch per kernel -----------------
 * No counterpart in the reference (one block per MPI job, src/block.cpp:191-215);
 * for serving many small independent boxes.  A pack advances every member
 * context by whole RK3 steps with one launch of each kernel for all members:
 * workgroup b of a packed grid runs member j's work for its own workgroup
 * b - off[j], so each member's result is bit-identical to nuc_step on that
 * member alone.  Members must be compatible (nuc_pack_compatible: same
 * schemes, splitting, model, viscous schemes, precision and line-length
 * class per direction, no forcing, lines of at most 128 cells); they keep
 * their own state, dt, residual and time.  The pack works on its own stream,
 * ordered after each member's earlier work and before its later work, so
 * nuc_sync/nuc_query/nuc_download_state on a member see the pack's steps.
 * use_graph: replay one captured hipGraph of a packed step.  Destroy a pack
 * before any of its members; 1 to 64 members. */
typedef struct nuc_pack nuc_pack;
int nuc_pack_compatible(const nuc_ctx *a, const nuc_ctx *b);
int nuc_pack_create(nuc_ctx *const *members, int n, int use_graph, nuc_pack **out);
int nuc_pack_step_async(nuc_pack *pack, int nsteps);
int nuc_pack_destroy(nuc_pack *pack);
/* the message of the last failure of pack (or of nuc_pack_create when pack is NULL) */
const char *nuc_pack_last_error(const nuc_pack *pack);
/* Global statistics of postproc::solveQ (src/postproc.cpp:736-929), FP64:
 * out[0..15] means of ux,uy,uz,vx,vy,vz,wx,wy,wz, omega_x,y,z, enstrophy, Q,
 * helicity, kinetic energy; out[16..24] u_rms, M_t, sqrt(<(du_i/dx_i)^2>/3),
 * mu_mean, Re_lambda, lambda, rho_mean, tau_t, skewness. */
int nuc_stats(nuc_ctx *ctx, double *out25);
/* The same 25 values followed by the dissipation rates the mixed path reports
 * its error on: out[25] = <mu |omega|^2>/Re (solenoidal), out[26] =
 * 4/3 <mu (div u)^2>/Re (dilatational), mu the Sutherland law of getMiu
 * (src/physicsModel.cpp:193-225).  Writes min(nout, NUC_NSTATS_EX) values. */
#define NUC_NSTATS_EX 27
int nuc_stats_ex(nuc_ctx *ctx, double *out, int nout);

/* ---- time-averaged statistics and TKE budget ------------------------------
 * postproc::solveAveraged0/1/2 (src/postproc.cpp:1186-2110) on the device, FP64
 * in both precision paths, field order of src/postproc.hpp:49-121.
 * nuc_average_begin(type): TimeAverageType 1 = primaries (10: rho u v w p T e
 *   alpha miu coeff) and gradients (21: d(u,v,w,p,T)/d(x,y,z), vorticity and its
 *   squares); 2 = + derived (28); 3 = + TKE fields (49) and the TKE budget (45:
 *   tke, production, transport, pressure diffusion p1 p2 p3, viscous diffusion v1
 *   v2 v3, convection, then the intermediate fields).  A new average starts from
 *   zero fields (TimeAverageStartStep 0).
 * nuc_average_update: one averaging step on the current state:
 *   f_aver = (f_aver*n + f)/(n + 1); with type 3 also the budget.
 * nuc_average_field: one field of group 0 primary, 1 grads, 2 derived, 3 TKE
 *   (all time-averaged) or 4 budget, N doubles.
 * nuc_average_means: volume means of every field of a group (fixed-order FP64
 *   reductions); out holds at least the group's field count. */
int nuc_average_begin(nuc_ctx *ctx, int type);
int nuc_average_update(nuc_ctx *ctx);
long long nuc_average_steps(const nuc_ctx *ctx);
int nuc_average_field(nuc_ctx *ctx, int group, int index, double *out);
int nuc_average_means(nuc_ctx *ctx, int group, double *out, int nout);
int nuc_average_end(nuc_ctx *ctx);
/* nuc_average_restore: start an average of `type` from saved fields instead of
 * zeros, as postproc's constructor does from an averaged-statistics file
 * (src/postproc.cpp:157-256): `fields` holds every field of the type's groups
 * in file order (primary, grads[, derived[, TKE, budget]]), N doubles each,
 * and `steps` the averaging steps they hold (averStep). */
int nuc_average_restore(nuc_ctx *ctx, int type, long long steps, const double *fields);

/* ---- the reference's files, per rank of the config's rank grid -----------
 * Rank (cx, cy, cz) is rank (cz*npy + cy)*npx + cx (the reference's Cartesian
 * communicator, src/MPICommunicator.cpp:26-33) and holds its block of the
 * global field, x fastest.  Files go to / come from directory `dir`.
 * nuc_write_restart / nuc_read_restart: block::outputQ_binary / inputQ_binary
 *   (src/block.cpp:609-671), NUC3d_step_<step>_ID_<rank>.bin = [double time]
 *   [5 fields]; q is the global state (5N doubles); *time gets rank 0's time.
 * nuc_write_tecplot: block::outputQ_tecplot (src/block.cpp:553-607),
 *   NUC3d_step_<step>_id_<rank>.dat: node coordinates (origin subtracted; NULL
 *   = 0) and rho u v w pressure temperature e sound_speed cell-centred.
 * nuc_average_write: postproc::OutputAveraged_bin (src/postproc.cpp:412-522),
 *   Averaged_step_<averStep>_id_<rank>.bin = [int averStep][groups of `type`];
 *   type must be the context's running averaging type.
 * nuc_average_read: the reader of src/postproc.cpp:157-256 for
 *   Averaged_step_<step>_ID_<rank>.bin (or _id_, as the writer names it), then
 *   nuc_average_restore. */
int nuc_write_restart(const nuc_config *cfg, const char *dir, int step, double time, const double *q);
int nuc_read_restart(const nuc_config *cfg, const char *dir, int step, double *time, double *q);
int nuc_write_tecplot(const nuc_config *cfg, const char *dir, int step, const double *origin, const double *q);
int nuc_average_write(nuc_ctx *ctx, const nuc_config *cfg, int type, const char *dir);
int nuc_average_read(nuc_ctx *ctx, const nuc_config *cfg, int type, const char *dir, int step);
/* nuc_average_write_tecplot: postproc::OutputAveraged_tec (src/postproc.cpp:523-720),
 *   Averaged_tec_step_<averStep>_id_<rank>.dat: title with the time, node
 *   coordinates (origin subtracted; NULL = 0) and the type's averaged fields
 *   cell-centred, 12 significant digits. */
int nuc_average_write_tecplot(nuc_ctx *ctx, const nuc_config *cfg, int type, const char *dir, double time,
                              const double *origin);
/* nuc_write_q_tecplot: postproc::OutputTemporal (src/postproc.cpp:346-405),
 *   Q_tec_step_<step>_id_<rank>.dat: title with the time, node coordinates and
 *   the 16 fields of nuc_q_fields of the current state, cell-centred. */
int nuc_write_q_tecplot(nuc_ctx *ctx, const nuc_config *cfg, const char *dir, int step, double time,
                        const double *origin);

/* ---- shell spectra ----------------------------------------------------------
 * postproc::solveFFT_* (src/postproc.cpp:931-1184) with fft::solveShell
 * (src/fft.cpp:384-572) of the current state; kmax = min(nx, ny, nz)/2 shells,
 * the forward transform e^{-i} divided by N.
 * nuc_spectra_kinetic: out[s*kmax + k], s = 0 total, 1 compressive,
 *   2 solenoidal, 3 left- and 4 right-helical kinetic energy of shell k (signed
 *   wave numbers; k = 0 counted in shell 1 as written).
 * nuc_spectrum_scalar: |f^|^2 of shell k for density, pressure, enstrophy
 *   0.5|omega|^2 or helicity omega.u, over the non-negative index octant as the
 *   reference's solveShell counts it.
 * Shell sums are FP64; their order within a workgroup is not fixed. */
enum nuc_spectrum_field { NUC_SPEC_DENSITY = 0, NUC_SPEC_PRESSURE = 1, NUC_SPEC_ENSTROPHY = 2, NUC_SPEC_HELICITY = 3 };
/* nuc_q_fields: the 16 per-point fields of postproc::solveQ
 * (src/postproc.cpp:736-835) of the current state into out (16N host doubles,
 * field after field): du/dx du/dy du/dz dv/dx ... dw/dz with the viscous
 * schemes, vorticity (3), enstrophy |omega|^2/2, Q, helicity omega.u, kinetic
 * energy rho|u|^2/2. */
int nuc_q_fields(nuc_ctx *ctx, double *out);
int nuc_spectra_kinetic(nuc_ctx *ctx, double *out, int kmax);
int nuc_spectrum_scalar(nuc_ctx *ctx, int field, double *out, int kmax);

/* ---- operator hooks (reference plug-in API, one Field per call) -----------
 * These mirror interoplation::interpolationInner/BoundaryL/BoundaryR
 * (src/fieldOperator.h:13-37), differential::differential* (:58-81) and
 * integration::integrationAll (:39-56) on host arrays, running the same
 * device code as the fused sweeps.  zones is a bit mask: 1 = inner
 * interfaces/cells (interpolationInner), 2 = left ghost-aware edge
 * (interpolationBoundaryL), 4 = right edge (interpolationBoundaryR); 7 = all.
 * gL may be NULL unless bit 2 is set, gR unless bit 4 is set. */
int nuc_op_reconstruct(int scheme, int upwind, int tilesize, const double *in, int n1, int n2, int n3,
                       const double *gL, const double *gR, int bf, double *out, int zones);
int nuc_op_diff(int vis_scheme, int tilesize, const double *in, double h, int n1, int n2, int n3,
                const double *gL, const double *gR, int bf, double *out, int zones);
int nuc_op_rk3(int stage, const double *rhs, const double *un, double *ui, double dt, long n);

/* ---- timing ----------------------------------------------------------------
 * With profiling on, every kernel launch of the context is bracketed by HIP
 * events on the context's stream; nuc_sync accumulates the elapsed times per
 * kernel slot (0 k_prim, 1 k_dt, 2 k_grad, 3/4/5 the x/y/z sweeps, 6 k_res).
 * nuc_profile(ctx, on) also clears the totals. */
int nuc_profile(nuc_ctx *ctx, int enable);
int nuc_profile_read(const nuc_ctx *ctx, int slot, double *total_ms, long *launches, const char **name);

/* ---- configuration -------------------------------------------------------
 * Fill cfg from a directory holding the reference's inp files
 * (inp/block.in, inp/Method.in, inp/PhysModel.in, inp/IOController.in) with the
 * reference's parsing rules, including the positional first six pairs of
 * PhysModel.in (src/physicsModel.cpp:72-77).  start/end steps and cfl come
 * from IOController.in (src/IOcontroller.cpp:12-110). */
int nuc_config_from_inp(const char *dir, nuc_config *cfg, int *start_step, int *end_step);

/* ---- host-side initial conditions (5N doubles into q) ---------------------
 * nuc_init_tgv: Taylor-Green vortex of initial_taylorgreen (src/block.cpp:359-395),
 *   cell centres placed per rank as getXYZ does (src/block.cpp:103-140).
 * nuc_init_hit: random-phase solenoidal velocity with
 *   E(k) ~ (k/k0)^4 exp(-2 (k/k0)^2) as initial_hit (src/block.cpp:397-551), but
 *   with phases drawn from mt19937_64(seed) instead of std::random_device and
 *   rescaled to the given u_rms; rho = 1, p = 1/(gamma Mach^2). */
int nuc_init_tgv(const nuc_config *cfg, double *q);
/* nuc_init_default: fluid at rest, rho = 1, p = 1/(gamma Mach^2), as
 *   initial_default (src/block.cpp:285-314) whose u_init, v_init, w_init
 *   PhysModel.in cannot set (always 0).
 * nuc_init_ivc: isentropic vortex of initial_ivc (src/block.cpp:316-357),
 *   centre (5, 5) in x-y, strength 0.5, carried by u = 0.5, p = rho^gamma. */
int nuc_init_default(const nuc_config *cfg, double *q);
int nuc_init_ivc(const nuc_config *cfg, double *q);
int nuc_init_hit(const nuc_config *cfg, double k0, double u_rms, unsigned long long seed, double *q);
/* nuc_init_shu_osher: Mach-3 shock / entropy-wave profile along x, measured
 *   from the box centre: (rho, u, p) = (3.857143, 2.629369, 10.33333) for
 *   x < -4, (1 + 0.2 sin 5x, 0, 1) elsewhere, plus a solenoidal velocity
 *   perturbation of rms pert * 2.629369 (nuc_init_hit's generator, k0 = 4,
 *   seed).  No reference counterpart: the reference's initial conditions
 *   (src/block.cpp:285-551) have no shock case, and its boundaries are periodic,
 *   so the seam at x = +-lx/2 is a second discontinuity.  With Mach =
 *   gamma^-1/2 the temperature T = gamma M^2 p / rho of src/physicsModel.cpp:342
 *   equals p / rho. */
int nuc_init_shu_osher(const nuc_config *cfg, double pert, unsigned long long seed, double *q);
/* The same fields generated on the device into ctx's state (a new step starts):
 * nuc_init_tgv_device, nuc_init_default_device and nuc_init_ivc_device are
 * nuc_init_tgv, nuc_init_default and nuc_init_ivc; nuc_init_hit_device synthesises
 * nuc_init_hit's seeded modes with a separable inverse DFT on the device (equal
 * to the host field up to rounding, ~1e-15 relative) -- for large boxes and for
 * jobs whose initial state should not cross PCIe. */
int nuc_init_tgv_device(nuc_ctx *ctx);
int nuc_init_default_device(nuc_ctx *ctx);
int nuc_init_ivc_device(nuc_ctx *ctx);
int nuc_init_hit_device(nuc_ctx *ctx, double k0, double u_rms, unsigned long long seed);

#ifdef __cplusplus
}
#endif
#endif
