// nuc3d_server.cpp -- the serving loop of include/nuc3d_serve.h.
//
// One worker thread owns every device call of the jobs it runs:
//   admit   : first-fit over the queue in arrival order, a job is admitted when
//             a slot is free and its nuc_footprint fits the remaining budget;
//             its context is created (stream-ordered allocations on its own
//             non-blocking stream, so running jobs are not stalled) and its
//             initial state uploaded;
//   advance : every admitted job whose stream is idle gets its next slice of
//             steps enqueued (hipGraph replays of one whole RK3 step), so the
//             jobs' kernels overlap on the device;
//   retire  : a job whose last slice is done is checked (finite dt and
//             residual), downloaded, its context destroyed and its bytes
//             returned to the budget.
//   pack    : idle jobs that can share kernels (nuc_pack_compatible) are
//             grouped, up to pack_cells cells per group, and a group of two or
//             more advances as one pack: one launch per kernel for all of its
//             members (nuc_pack_step_async), each member's result unchanged.
//             Packs are kept while their membership recurs (the members of a
//             pack finish their slices together, so the same group usually
//             forms again) and dropped when a member retires.
// Slice length (steps per turn) is sized per job from a model of its own step
// time -- a fixed launch cost plus its cells at the device's small-box
// throughput -- aiming at about SLICE_TARGET_MS per turn, with at least
// SLICE_MIN steps: long enough that the host's per-turn work (query, sync,
// enqueue) stays off the critical path, short enough that a large box does
// not hold its slot for long.  Wall-clock slice times are not used: with many
// jobs in flight they measure the queueing behind other jobs, not the job.
// A pack takes the shortest slice of its members.
#include "nuc3d_serve.h"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace {

constexpr double SLICE_TARGET_MS = 8.0;
constexpr int SLICE_MIN = 4;
constexpr int SLICE_MAX = 32;
// modelled step time of a box alone: launches of one RK3 step, plus three
// stages of its cells at a small-box throughput of 2.5e9 cell-updates/s
inline double step_ms_model(long long cells) { return 0.25 + 3.0 * (double)cells / 2.5e9 * 1e3; }
constexpr size_t MAX_PACKS = 16;

struct Job {
    long long id = 0;
    nuc_config cfg{};
    std::vector<double> q;     // initial state, then the final state
    int nsteps = 0, done = 0, inflight = 0, slice = 1;
    long long bytes = 0;
    nuc_ctx *ctx = nullptr;
    nuc_job_result r{};
    std::string err;
    long long cells = 0;
    bool packable = false;
    // initial state generated on the device at admission instead of q
    int ic = NUC_IC_STATE;
    double ic_k0 = 0.0, ic_urms = 0.0;
    unsigned long long ic_seed = 0;
};

struct PackRec {
    std::vector<long long> ids;   // member job ids, in pack order
    nuc_pack *pack = nullptr;
};

}  // namespace

struct nuc_server {
    nuc_server_config cfg{};
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<Job>> queue;
    std::vector<std::shared_ptr<Job>> active;     // owned by the worker
    std::map<long long, std::shared_ptr<Job>> jobs;
    long long next_id = 1;
    long long bytes_in_use = 0, peak_bytes = 0, completed = 0, failed = 0, peak_active = 0;
    long long n_active = 0;
    long long single_slices = 0, pack_slices = 0;
    std::vector<PackRec> packs;                   // owned by the worker
    bool stop = false;
    std::thread worker;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();

    double now() const
    {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    void run();
    void finish(const std::shared_ptr<Job> &j, int state, int error, const std::string &msg);
    void drop_packs_of(long long id);
    void launch_idle(std::vector<std::shared_ptr<Job>> &idle);
    bool start_slice(const std::shared_ptr<Job> &j);
};

void nuc_server::drop_packs_of(long long id)
{
    for (size_t i = 0; i < packs.size();) {
        if (std::find(packs[i].ids.begin(), packs[i].ids.end(), id) != packs[i].ids.end()) {
            nuc_pack_destroy(packs[i].pack);
            packs.erase(packs.begin() + i);
        } else {
            ++i;
        }
    }
}

void nuc_server::finish(const std::shared_ptr<Job> &j, int state, int error, const std::string &msg)
{
    drop_packs_of(j->id);   // a pack must go before its members
    if (j->ctx) {
        nuc_destroy(j->ctx);
        j->ctx = nullptr;
    }
    std::lock_guard<std::mutex> lk(mu);
    j->r.state = state;
    j->r.error = error;
    j->r.done_s = now();
    j->err = msg;
    if (state == NUC_JOB_FAILED) j->q.clear();
    bytes_in_use -= j->bytes;
    n_active--;
    if (state == NUC_JOB_DONE) completed++;
    else failed++;
    cv_done.notify_all();
}

void nuc_server::run()
{
    for (;;) {
        std::vector<std::shared_ptr<Job>> admit;
        {
            std::unique_lock<std::mutex> lk(mu);
            if (stop) return;
            // admission: first fit in arrival order
            for (auto it = queue.begin(); it != queue.end();) {
                if ((long long)active.size() + (long long)admit.size() >= cfg.max_active) break;
                const auto &j = *it;
                if (bytes_in_use + j->bytes <= cfg.hbm_budget) {
                    bytes_in_use += j->bytes;
                    peak_bytes = std::max(peak_bytes, bytes_in_use);
                    n_active++;
                    peak_active = std::max(peak_active, n_active);
                    j->r.state = NUC_JOB_RUNNING;
                    j->r.admit_s = now();
                    admit.push_back(j);
                    it = queue.erase(it);
                } else {
                    ++it;
                }
            }
        }
        bool progressed = !admit.empty();
        for (auto &j : admit) {
            int rc = nuc_create(&j->cfg, &j->ctx);
            if (rc != NUC_OK) {
                j->ctx = nullptr;
                finish(j, NUC_JOB_FAILED, rc, "nuc_create failed");
                continue;
            }
            if (cfg.use_graphs) nuc_set_graph(j->ctx, 1);
            j->cells = (long long)j->cfg.nx * j->cfg.ny * j->cfg.nz;
            j->packable = nuc_pack_compatible(j->ctx, j->ctx) != 0;
            if (j->ic == NUC_IC_TGV) rc = nuc_init_tgv_device(j->ctx);
            else if (j->ic == NUC_IC_HIT) rc = nuc_init_hit_device(j->ctx, j->ic_k0, j->ic_urms, j->ic_seed);
            else if (j->ic == NUC_IC_IVC) rc = nuc_init_ivc_device(j->ctx);
            else if (j->ic == NUC_IC_DEFAULT) rc = nuc_init_default_device(j->ctx);
            else rc = nuc_upload_state(j->ctx, j->q.data());
            if (j->ic != NUC_IC_STATE) j->q.assign((size_t)5 * j->cfg.nx * j->cfg.ny * j->cfg.nz, 0.0);
            if (rc == NUC_OK && cfg.autotune) rc = nuc_autotune(j->ctx);
            if (rc != NUC_OK) {
                finish(j, NUC_JOB_FAILED, rc, nuc_last_error(j->ctx));
                continue;
            }
            if (cfg.steps_per_slice > 0) {
                j->slice = cfg.steps_per_slice;
            } else {
                int k = (int)(SLICE_TARGET_MS / step_ms_model((long long)j->cfg.nx * j->cfg.ny * j->cfg.nz));
                j->slice = std::min(SLICE_MAX, std::max(SLICE_MIN, k));
            }
            active.push_back(j);
        }
        std::vector<std::shared_ptr<Job>> idle;
        for (size_t a = 0; a < active.size();) {
            auto j = active[a];
            int q = nuc_query(j->ctx);
            if (q == NUC_EBUSY) { ++a; continue; }
            bool retire = false;
            if (q != NUC_OK) {
                finish(j, NUC_JOB_FAILED, q, nuc_last_error(j->ctx));
                retire = true;
            } else {
                if (j->inflight > 0) {
                    nuc_step_info info;
                    int rc = nuc_sync(j->ctx, &info);
                    if (rc != NUC_OK) {
                        finish(j, NUC_JOB_FAILED, rc, nuc_last_error(j->ctx));
                        retire = true;
                    } else {
                        j->r.info = info;
                        j->inflight = 0;
                    }
                }
                if (!retire && j->done >= j->nsteps) {
                    int rc = nuc_download_state(j->ctx, j->q.data());
                    if (rc == NUC_OK) finish(j, NUC_JOB_DONE, NUC_OK, "");
                    else finish(j, NUC_JOB_FAILED, rc, nuc_last_error(j->ctx));
                    retire = true;
                } else if (!retire) {
                    idle.push_back(j);
                }
            }
            progressed = true;
            if (retire) active.erase(active.begin() + a);
            else ++a;
        }
        launch_idle(idle);
        if (!progressed) {
            std::unique_lock<std::mutex> lk(mu);
            if (stop) return;
            // idle: every admitted job is busy on the device, or nothing to do
            cv_work.wait_for(lk, std::chrono::microseconds(active.empty() ? 2000 : 50));
        }
    }
}

// one slice of a job on its own stream; false when the job failed (retired)
bool nuc_server::start_slice(const std::shared_ptr<Job> &j)
{
    int k = std::min(j->slice, j->nsteps - j->done);
    int rc = nuc_step_async(j->ctx, k);
    if (rc != NUC_OK) {
        finish(j, NUC_JOB_FAILED, rc, nuc_last_error(j->ctx));
        active.erase(std::find(active.begin(), active.end(), j));
        return false;
    }
    j->done += k;
    j->inflight = k;
    single_slices++;
    return true;
}

void nuc_server::launch_idle(std::vector<std::shared_ptr<Job>> &idle)
{
    std::vector<bool> used(idle.size(), false);
    for (size_t a = 0; a < idle.size(); a++) {
        if (used[a]) continue;
        used[a] = true;
        std::vector<std::shared_ptr<Job>> grp{idle[a]};
        if (cfg.pack_cells > 0 && idle[a]->packable) {
            long long cells = idle[a]->cells;
            for (size_t b = a + 1; b < idle.size() && grp.size() < 64; b++) {
                if (used[b] || !idle[b]->packable || cells + idle[b]->cells > cfg.pack_cells) continue;
                if (!nuc_pack_compatible(idle[a]->ctx, idle[b]->ctx)) continue;
                used[b] = true;
                cells += idle[b]->cells;
                grp.push_back(idle[b]);
            }
        }
        if (grp.size() == 1) {
            start_slice(grp[0]);
            continue;
        }
        std::vector<long long> ids;
        for (auto &j : grp) ids.push_back(j->id);
        PackRec *rec = nullptr;
        for (auto &p : packs)
            if (p.ids == ids) { rec = &p; break; }
        if (!rec) {
            // bounded cache: drop a pack none of whose members has work in flight
            if (packs.size() >= MAX_PACKS) {
                for (size_t i = 0; i < packs.size(); i++) {
                    bool busy = false;
                    for (long long id : packs[i].ids)
                        for (auto &j : active)
                            if (j->id == id && j->inflight > 0) busy = true;
                    if (!busy) {
                        nuc_pack_destroy(packs[i].pack);
                        packs.erase(packs.begin() + i);
                        break;
                    }
                }
            }
            std::vector<nuc_ctx *> m;
            for (auto &j : grp) m.push_back(j->ctx);
            nuc_pack *pk = nullptr;
            if (nuc_pack_create(m.data(), (int)m.size(), cfg.use_graphs, &pk) != NUC_OK) {
                for (auto &j : grp) start_slice(j);   // run them one by one instead
                continue;
            }
            packs.push_back(PackRec{ids, pk});
            rec = &packs.back();
        }
        int k = SLICE_MAX;
        for (auto &j : grp) k = std::min(k, std::min(j->slice, j->nsteps - j->done));
        if (nuc_pack_step_async(rec->pack, k) != NUC_OK) {
            drop_packs_of(grp[0]->id);
            for (auto &j : grp) start_slice(j);
            continue;
        }
        for (auto &j : grp) {
            j->done += k;
            j->inflight = k;
        }
        pack_slices++;
    }
}

extern "C" int nuc_server_create(const nuc_server_config *cfg, nuc_server **out)
{
    if (!cfg || !out || cfg->max_active < 1 || cfg->hbm_budget <= 0) return NUC_EINVAL;
    nuc_server *s = new nuc_server();
    s->cfg = *cfg;
    s->worker = std::thread([s] { s->run(); });
    *out = s;
    return NUC_OK;
}

extern "C" int nuc_server_destroy(nuc_server *s)
{
    if (!s) return NUC_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->worker.joinable()) s->worker.join();
    for (auto &p : s->packs) nuc_pack_destroy(p.pack);
    s->packs.clear();
    for (auto &j : s->active)
        if (j->ctx) nuc_destroy(j->ctx);
    delete s;
    return NUC_OK;
}

static int submit_job(nuc_server *s, const std::shared_ptr<Job> &j, long long *job_id);

extern "C" int nuc_server_submit(nuc_server *s, const nuc_config *cfg, const double *q0, int nsteps,
                                 long long *job_id)
{
    if (!s || !cfg || !q0 || nsteps < 0 || !job_id) return NUC_EINVAL;
    auto j = std::make_shared<Job>();
    j->cfg = *cfg;
    size_t n = (size_t)5 * cfg->nx * cfg->ny * cfg->nz;
    if (cfg->nx < 1 || cfg->ny < 1 || cfg->nz < 1) return NUC_EINVAL;
    j->q.assign(q0, q0 + n);
    j->nsteps = nsteps;
    return submit_job(s, j, job_id);
}

extern "C" int nuc_server_submit_ic(nuc_server *s, const nuc_config *cfg, int ic, double k0, double u_rms,
                                    unsigned long long seed, int nsteps, long long *job_id)
{
    if (!s || !cfg || nsteps < 0 || !job_id || ic < NUC_IC_TGV || ic > NUC_IC_DEFAULT) return NUC_EINVAL;
    if (cfg->nx < 1 || cfg->ny < 1 || cfg->nz < 1) return NUC_EINVAL;
    if (ic == NUC_IC_HIT && (!(k0 > 0) || !(u_rms >= 0))) return NUC_EINVAL;
    auto j = std::make_shared<Job>();
    j->cfg = *cfg;
    j->ic = ic;
    j->ic_k0 = k0;
    j->ic_urms = u_rms;
    j->ic_seed = seed;
    j->nsteps = nsteps;
    return submit_job(s, j, job_id);
}

static int submit_job(nuc_server *s, const std::shared_ptr<Job> &j, long long *job_id)
{
    const nuc_config *cfg = &j->cfg;
    j->bytes = nuc_footprint(cfg);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (j->bytes > s->cfg.hbm_budget) return NUC_ENOMEM;
        j->id = s->next_id++;
        j->r.state = NUC_JOB_QUEUED;
        j->r.submit_s = s->now();
        s->queue.push_back(j);
        s->jobs[j->id] = j;
        *job_id = j->id;
    }
    s->cv_work.notify_all();
    return NUC_OK;
}

extern "C" int nuc_server_result(nuc_server *s, long long id, double *q_out, nuc_job_result *r)
{
    if (!s) return NUC_EINVAL;
    std::shared_ptr<Job> j;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->jobs.find(id);
        if (it == s->jobs.end()) return NUC_EINVAL;
        j = it->second;
        if (j->r.state != NUC_JOB_DONE && j->r.state != NUC_JOB_FAILED) return NUC_EBUSY;
        s->jobs.erase(it);
    }
    if (r) *r = j->r;
    if (q_out && j->r.state == NUC_JOB_DONE) memcpy(q_out, j->q.data(), j->q.size() * sizeof(double));
    return NUC_OK;
}

extern "C" int nuc_server_drain(nuc_server *s, double timeout_s)
{
    if (!s) return NUC_EINVAL;
    std::unique_lock<std::mutex> lk(s->mu);
    auto idle = [s] { return s->queue.empty() && s->n_active == 0; };
    if (timeout_s < 0) {
        s->cv_done.wait(lk, idle);
        return NUC_OK;
    }
    return s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), idle) ? NUC_OK : NUC_EBUSY;
}

extern "C" int nuc_server_stats_get(nuc_server *s, nuc_server_stats *st)
{
    if (!s || !st) return NUC_EINVAL;
    std::lock_guard<std::mutex> lk(s->mu);
    st->completed = s->completed;
    st->failed = s->failed;
    st->queued = (long long)s->queue.size();
    st->active = s->n_active;
    st->peak_active = s->peak_active;
    st->bytes_in_use = s->bytes_in_use;
    st->peak_bytes = s->peak_bytes;
    st->single_slices = s->single_slices;
    st->pack_slices = s->pack_slices;
    return NUC_OK;
}

extern "C" double nuc_server_now(nuc_server *s) { return s ? s->now() : 0.0; }

extern "C" const char *nuc_server_job_error(nuc_server *s, long long id)
{
    // copied under the lock into this thread's buffer: a concurrent
    // nuc_server_result may free the job right after the lock is released
    static thread_local std::string msg;
    if (!s) return "null server";
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(id);
    msg = (it == s->jobs.end()) ? std::string("unknown job") : it->second->err;
    return msg.c_str();
}
