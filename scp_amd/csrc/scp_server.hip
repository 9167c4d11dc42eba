// scp_server.hip — a request loop over device-resident grids (scp_server_*).
//
// Jobs arrive as (parameters, Q-balls, seed, steps). Each scheduling round
//   1. admits queued jobs in arrival order while their device bytes
//      (scp_grid_bytes) fit the HBM budget: the grid is created, seeded on the
//      device (scp_grid_seed_balls), Gauss-projected if asked, and joins the
//      batch of its class;
//   2. advances every active job by min(quantum, steps left) in one batched
//      chain per class (scp_batch_step): jobs admitted this round share the
//      launches of jobs admitted earlier (continuous batching);
//   3. retires jobs that reached their step count; their grids stay resident
//      until the caller takes the result (scp_server_take), which frees the
//      memory for the next admissions.
// The batched step is bitwise the single-grid step and the FP32C step does not
// depend on how a run is split into chains, so a job's result does not depend
// on what it shared the GPU with.
#include "scp_grid.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <vector>

using namespace scp;
using namespace scp_internal;

namespace {

struct Job {
    int64_t id = 0;
    scp_params p{};
    std::vector<scp_qball> balls;
    double theta_sigma = 0, link_sigma = 0;
    uint64_t seed = 0;
    int nsteps = 0, project = 0;
    int state = SCP_JOB_QUEUED;
    int64_t done_steps = 0, bytes = 0;
    double t_submit = 0, t_admit = 0, t_done = 0;
    scp_grid *g = nullptr;
};

struct ClassBatch {
    scp_batch *b = nullptr;
    std::vector<Job *> members;   // same order as the batch's members
};

int class_key(const scp_params &p) {
    const int step_mode = (p.mode == 2 && p.g == 0.0) ? 1 : p.mode;
    return step_mode | (p.precision << 4) | ((p.precision == SCP_PREC_F32C ? p.lo_words : 0) << 8);
}

}  // namespace

struct scp_server {
    int64_t budget = 0, in_use = 0;
    int quantum = 1, max_active = 0;
    std::deque<Job *> queue;
    std::map<int64_t, Job *> jobs;
    std::map<int, ClassBatch> classes;
    int64_t next_id = 1, site_updates = 0;
    double busy = 0;
    std::chrono::steady_clock::time_point t0;
    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
};

static void release_job_grid(scp_server *s, Job *j) {
    if (j->g) {
        scp_grid_destroy(j->g);   // also leaves its batch
        j->g = nullptr;
        s->in_use -= j->bytes;
    }
}

extern "C" int scp_server_create(int64_t hbm_budget, int quantum, int max_active, scp_server **out) {
    if (!out) return set_err(-1, "null out");
    *out = nullptr;
    if (hbm_budget <= 0) return set_err(-1, "hbm_budget must be > 0");
    if (quantum < 1) return set_err(-1, "quantum=%d must be >= 1", quantum);
    if (max_active < 0) return set_err(-1, "max_active=%d must be >= 0", max_active);
    scp_server *s = new scp_server;
    s->budget = hbm_budget;
    s->quantum = quantum;
    s->max_active = max_active;
    s->t0 = std::chrono::steady_clock::now();
    *out = s;
    return 0;
}

extern "C" int scp_server_destroy(scp_server *s) {
    if (!s) return 0;
    for (auto &kv : s->jobs) {
        release_job_grid(s, kv.second);
        delete kv.second;
    }
    for (auto &kv : s->classes) scp_batch_destroy(kv.second.b);
    delete s;
    return 0;
}

extern "C" double scp_server_now(const scp_server *s) { return s ? s->now() : 0.0; }

extern "C" int scp_server_submit(scp_server *s, const scp_job *job, int64_t *id) {
    if (!s || !job || !id) return set_err(-1, "null argument");
    *id = 0;
    if (job->params.mode == 0) return set_err(-1, "jobs run the complex or gauged step");
    if (job->nsteps < 1) return set_err(-1, "nsteps=%d must be >= 1", job->nsteps);
    if (job->nballs < 0 || (job->nballs > 0 && !job->balls)) return set_err(-1, "bad ball list");
    if (job->params.step_kernel != 0) return set_err(-1, "jobs run the LDS-staged march (step_kernel 0)");
    if (job->params.bc_type == 1) return set_err(-1, "jobs run the periodic or sponge boundary (bc_type 0 or 2)");
    const int64_t bytes = scp_grid_bytes(&job->params);
    if (bytes < 0) return -1;   // message set by the parameter check
    if (bytes > s->budget)
        return set_err(-1, "job needs %lld bytes of HBM, more than the budget %lld", (long long)bytes,
                       (long long)s->budget);
    Job *j = new Job;
    j->id = s->next_id++;
    j->p = job->params;
    j->balls.assign(job->balls, job->balls + job->nballs);
    j->theta_sigma = job->theta_sigma;
    j->link_sigma = job->link_sigma;
    j->seed = job->seed;
    j->nsteps = job->nsteps;
    j->project = job->gauss_project;
    j->bytes = bytes;
    j->t_submit = s->now();
    s->queue.push_back(j);
    s->jobs[j->id] = j;
    *id = j->id;
    return 0;
}

// create, seed, project and enrol one job; on failure the job is marked failed
static int admit(scp_server *s, Job *j) {
    int rc = scp_grid_create(&j->p, &j->g);
    if (!rc) {
        s->in_use += j->bytes;
        rc = scp_grid_seed_balls(j->g, j->balls.data(), (int)j->balls.size(), j->theta_sigma, j->link_sigma, j->seed);
    }
    if (!rc && j->project && j->p.mode == 2 && j->p.g != 0.0) {
        int it = 0;
        double gm = 0;
        rc = scp_grid_gauss_project(j->g, &it, &gm);
    }
    ClassBatch &cb = s->classes[class_key(j->p)];
    if (!rc && !cb.b) rc = scp_batch_create(&cb.b);
    if (!rc) rc = scp_batch_add(cb.b, j->g);
    if (rc) {
        release_job_grid(s, j);
        j->state = SCP_JOB_FAILED;
        j->t_done = s->now();
        return rc;
    }
    cb.members.push_back(j);
    j->state = SCP_JOB_ACTIVE;
    j->t_admit = s->now();
    return 0;
}

extern "C" int scp_server_round(scp_server *s, int *finished) {
    if (!s) return set_err(-1, "null server");
    if (finished) *finished = 0;
    // 1. admissions, arrival order, while memory (and the slot limit) allow
    int active = 0;
    for (auto &kv : s->classes) active += (int)kv.second.members.size();
    while (!s->queue.empty()) {
        Job *j = s->queue.front();
        if (s->max_active && active >= s->max_active) break;
        if (s->in_use + j->bytes > s->budget) break;
        s->queue.pop_front();
        if (admit(s, j) == 0) active++;
    }
    // 2. one chain per class
    int done = 0;
    for (auto &kv : s->classes) {
        ClassBatch &cb = kv.second;
        if (cb.members.empty()) continue;
        std::vector<int> counts(cb.members.size());
        int64_t sites = 0;
        for (size_t m = 0; m < cb.members.size(); m++) {
            const Job *j = cb.members[m];
            counts[m] = (int)std::min<int64_t>(s->quantum, j->nsteps - j->done_steps);
            sites += (int64_t)counts[m] * j->p.N * j->p.N * j->p.N;
        }
        const double t = s->now();
        int rc = scp_batch_step(cb.b, counts.data(), (int)counts.size());
        s->busy += s->now() - t;
        if (rc) return rc;
        s->site_updates += sites;
        // 3. retire jobs that reached their step count (their grids stay until taken)
        std::vector<Job *> keep;
        for (size_t m = 0; m < cb.members.size(); m++) {
            Job *j = cb.members[m];
            j->done_steps += counts[m];
            if (j->done_steps >= j->nsteps) {
                scp_batch_remove(cb.b, j->g);
                j->state = SCP_JOB_DONE;
                j->t_done = s->now();
                done++;
            } else {
                keep.push_back(j);
            }
        }
        cb.members.swap(keep);
    }
    if (finished) *finished = done;
    return 0;
}

extern "C" int scp_server_job(const scp_server *s, int64_t id, scp_job_info *info) {
    if (!s || !info) return set_err(-1, "null argument");
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return set_err(-1, "no job %lld", (long long)id);
    const Job *j = it->second;
    memset(info, 0, sizeof *info);
    info->state = j->state;
    info->nsteps = j->nsteps;
    info->steps_done = j->done_steps;
    info->bytes = j->bytes;
    info->t_submit = j->t_submit;
    info->t_admit = j->t_admit;
    info->t_done = j->t_done;
    return 0;
}

extern "C" int scp_server_take(scp_server *s, int64_t id, double *const *cols, int ncols, double *diag, int ndiag) {
    if (!s) return set_err(-1, "null server");
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return set_err(-1, "no job %lld", (long long)id);
    Job *j = it->second;
    if (j->state == SCP_JOB_FAILED) {
        s->jobs.erase(it);
        delete j;
        return set_err(-1, "job %lld failed at admission", (long long)id);
    }
    if (j->state != SCP_JOB_DONE) return set_err(-1, "job %lld is not done", (long long)id);
    int rc = 0;
    if (cols) rc = scp_grid_download(j->g, cols, ncols);
    if (!rc && diag) rc = scp_grid_diag(j->g, diag, ndiag);
    if (rc) return rc;
    release_job_grid(s, j);
    s->jobs.erase(it);
    delete j;
    return 0;
}

extern "C" int scp_server_stats_get(const scp_server *s, scp_server_stats *st) {
    if (!s || !st) return set_err(-1, "null argument");
    memset(st, 0, sizeof *st);
    for (auto &kv : s->jobs) {
        const int state = kv.second->state;
        if (state == SCP_JOB_QUEUED) st->queued++;
        else if (state == SCP_JOB_ACTIVE) st->active++;
        else if (state == SCP_JOB_DONE) st->done++;
        else st->failed++;
    }
    st->bytes_in_use = s->in_use;
    st->site_updates = s->site_updates;
    st->busy_seconds = s->busy;
    st->now = s->now();
    return 0;
}
