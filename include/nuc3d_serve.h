/*
 * nuc3d_serve.h -- C ABI of the serving loop: many independent box
 * simulations (uneven grids, Mach, Re, schemes) advanced concurrently on one
 * MI355X with their state resident in HBM.
 *
 * The reference has no counterpart: it runs one box per MPI job through
 * block::solve (src/block.cpp:191-215).  A job here is exactly that solve --
 * nsteps TVD-RK3 steps of one nuc_config from an initial state -- and its
 * result is bit-identical to the same job run alone on a plain context
 * (nuc_create + nuc_step), because every job owns its buffers and the kernels
 * are deterministic.
 *
 * Continuous batching: a worker thread admits queued jobs whenever a slot and
 * HBM budget (nuc_footprint of the job) are free, keeps every admitted job's
 * work in flight on its own stream as hipGraph replays of whole RK3 steps, and
 * retires finished jobs, so jobs that arrive while others run join at once.
 * Small compatible jobs that are idle together are advanced as one pack (one
 * launch per kernel for all of them, nuc_pack_*), which leaves each result
 * unchanged.
 * Calls are thread-safe.  Errors of one job never stop the loop.
 */
#ifndef NUC3D_SERVE_H
#define NUC3D_SERVE_H
#include "nuc3d_gpu.h"
#ifdef __cplusplus
extern "C" {
#endif

typedef struct nuc_server_config {
    long long hbm_budget;   /* bytes of device memory the admitted jobs may hold together */
    int max_active;         /* jobs in flight at once (each has its own stream) */
    int steps_per_slice;    /* RK3 steps enqueued per job per scheduling turn (0: tuned per job) */
    int use_graphs;         /* replay captured hipGraph steps (1) or launch kernels (0) */
    long long pack_cells;   /* idle compatible jobs (nuc_pack_compatible) of together at most this
                               many cells are advanced as one pack (nuc_pack_*); 0: no packing */
    int autotune;           /* pick each job's tile shapes with nuc_autotune (cached per shape) */
} nuc_server_config;

enum nuc_job_state { NUC_JOB_QUEUED = 0, NUC_JOB_RUNNING = 1, NUC_JOB_DONE = 2, NUC_JOB_FAILED = 3 };

typedef struct nuc_job_result {
    nuc_step_info info;     /* of the job's last step */
    double submit_s;        /* times on the server's monotonic clock, seconds */
    double admit_s;
    double done_s;
    int state;              /* enum nuc_job_state */
    int error;              /* nuc_status of a failed job */
} nuc_job_result;

typedef struct nuc_server_stats {
    long long completed, failed, queued, active, peak_active;
    long long bytes_in_use, peak_bytes;
    long long single_slices;   /* slices enqueued for one job on its own stream */
    long long pack_slices;     /* slices enqueued for a pack of jobs */
} nuc_server_stats;

typedef struct nuc_server nuc_server;

int nuc_server_create(const nuc_server_config *cfg, nuc_server **out);
/* finishes nothing: in-flight and queued jobs are dropped */
int nuc_server_destroy(nuc_server *srv);
/* copies q0 (5*nx*ny*nz doubles); job_id receives the job's id */
int nuc_server_submit(nuc_server *srv, const nuc_config *cfg, const double *q0, int nsteps, long long *job_id);
/* a job whose initial state is generated on the device at admission (no host
 * state crosses PCIe): NUC_IC_TGV (nuc_init_tgv_device), NUC_IC_HIT
 * (nuc_init_hit_device with k0, u_rms, seed), NUC_IC_IVC (nuc_init_ivc_device)
 * or NUC_IC_DEFAULT (nuc_init_default_device) */
enum nuc_initial_condition { NUC_IC_STATE = 0, NUC_IC_TGV = 1, NUC_IC_HIT = 2, NUC_IC_IVC = 3, NUC_IC_DEFAULT = 4 };
int nuc_server_submit_ic(nuc_server *srv, const nuc_config *cfg, int ic, double k0, double u_rms,
                         unsigned long long seed, int nsteps, long long *job_id);
/* NUC_OK with r filled (state DONE or FAILED) and, if q_out is not NULL and the
 * job is done, its final state copied out; NUC_EBUSY while queued or running.
 * A job's result can be fetched once; afterwards its id is unknown (NUC_EINVAL). */
int nuc_server_result(nuc_server *srv, long long job_id, double *q_out, nuc_job_result *r);
/* waits until no job is queued or running, or timeout_s passes (NUC_EBUSY) */
int nuc_server_drain(nuc_server *srv, double timeout_s);
int nuc_server_stats_get(nuc_server *srv, nuc_server_stats *st);
/* seconds on the server's clock (the clock of submit_s/admit_s/done_s) */
double nuc_server_now(nuc_server *srv);
const char *nuc_server_job_error(nuc_server *srv, long long job_id);

#ifdef __cplusplus
}
#endif
#endif
