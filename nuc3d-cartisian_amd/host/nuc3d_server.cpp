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
// Slice length is tuned per job from the observed duration of its own slices,
// aiming at about SLICE_TARGET_MS of device work per turn: long enough to keep
// the host off the critical path for small boxes, short enough that a large
// box does not hold its slot's turn for long.
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

constexpr double SLICE_TARGET_MS = 4.0;
constexpr int SLICE_MAX = 64;

struct Job {
    long long id = 0;
    nuc_config cfg{};
    std::vector<double> q;     // initial state, then the final state
    int nsteps = 0, done = 0, inflight = 0, slice = 1;
    long long bytes = 0;
    nuc_ctx *ctx = nullptr;
    nuc_job_result r{};
    std::string err;
    double slice_t0 = 0.0;
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
    bool stop = false;
    std::thread worker;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();

    double now() const
    {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    void run();
    void finish(const std::shared_ptr<Job> &j, int state, int error, const std::string &msg);
};

void nuc_server::finish(const std::shared_ptr<Job> &j, int state, int error, const std::string &msg)
{
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
            rc = nuc_upload_state(j->ctx, j->q.data());
            if (rc != NUC_OK) {
                finish(j, NUC_JOB_FAILED, rc, nuc_last_error(j->ctx));
                continue;
            }
            j->slice = cfg.steps_per_slice > 0 ? cfg.steps_per_slice : 1;
            active.push_back(j);
        }
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
                        if (cfg.steps_per_slice <= 0) {
                            double per_step_ms = (now() - j->slice_t0) * 1e3 / j->inflight;
                            int s = (int)(SLICE_TARGET_MS / std::max(per_step_ms, 1e-3));
                            j->slice = std::min(SLICE_MAX, std::max(1, s));
                        }
                        j->inflight = 0;
                    }
                }
                if (!retire && j->done >= j->nsteps) {
                    int rc = nuc_download_state(j->ctx, j->q.data());
                    if (rc == NUC_OK) finish(j, NUC_JOB_DONE, NUC_OK, "");
                    else finish(j, NUC_JOB_FAILED, rc, nuc_last_error(j->ctx));
                    retire = true;
                } else if (!retire) {
                    int k = std::min(j->slice, j->nsteps - j->done);
                    j->slice_t0 = now();
                    int rc = nuc_step_async(j->ctx, k);
                    if (rc != NUC_OK) {
                        finish(j, NUC_JOB_FAILED, rc, nuc_last_error(j->ctx));
                        retire = true;
                    } else {
                        j->done += k;
                        j->inflight = k;
                    }
                }
            }
            progressed = true;
            if (retire) active.erase(active.begin() + a);
            else ++a;
        }
        if (!progressed) {
            std::unique_lock<std::mutex> lk(mu);
            if (stop) return;
            // idle: every admitted job is busy on the device, or nothing to do
            cv_work.wait_for(lk, std::chrono::microseconds(active.empty() ? 2000 : 50));
        }
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
    for (auto &j : s->active)
        if (j->ctx) nuc_destroy(j->ctx);
    delete s;
    return NUC_OK;
}

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
    return NUC_OK;
}

extern "C" double nuc_server_now(nuc_server *s) { return s ? s->now() : 0.0; }

extern "C" const char *nuc_server_job_error(nuc_server *s, long long id)
{
    if (!s) return "null server";
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(id);
    return it == s->jobs.end() ? "unknown job" : it->second->err.c_str();
}
