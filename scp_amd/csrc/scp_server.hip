// scp_server.hip — a request loop over device-resident grids (scp_server_*).
//
// Jobs arrive as (parameters, Q-balls, seed, steps). Each scheduling round
//   1. admits queued jobs in arrival order while their device bytes fit the
//      HBM budget: the grid is created, seeded on the device
//      (scp_grid_seed_balls), Gauss-projected if asked, and joins the batch of
//      its class. A queued job that does not fit is passed over by later jobs
//      that do for at most `aging` rounds; after that it holds the line until
//      memory frees up for it;
//   2. advances every active job by min(quantum, steps left) in one batched
//      chain per class (scp_batch_step): jobs admitted this round share the
//      launches of jobs admitted earlier (continuous batching);
//   3. retires jobs that reached their step count; their grids stay resident
//      until the caller takes the result (scp_server_take), which frees the
//      memory for the next admissions.
// Device memory comes from a pool of large slabs (first use grows it; freed
// ranges are reused, coalesced), so admissions and retirements do not call
// hipMalloc / hipFree once the pool has grown. A job holds its FP64 emit
// buffers only while it needs them: at admission (the device seed and the
// projection work on them) and from its last chain on (the result lives
// there); in between its grid holds the step state alone (an FP32C gauged job
// with int8 position words: 212 instead of 452 bytes per site), and after its
// last step the step state is released.
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

// Device memory pool: slabs from hipMalloc, carved best-fit in 256-byte units;
// a release returns the range to its slab's free map, merged with its neighbours.
class DevPool {
  public:
    ~DevPool() {
        for (auto &kv : slabs_) hipFree(kv.first);
    }
    hipError_t alloc(void **p, size_t bytes) {
        *p = nullptr;
        bytes = (std::max<size_t>(bytes, 1) + 255) / 256 * 256;
        Slab *best = nullptr;
        size_t best_off = 0, best_size = SIZE_MAX;
        for (auto &kv : slabs_)
            for (auto &fr : kv.second.free)
                if (fr.second >= bytes && fr.second < best_size) {
                    best = &kv.second;
                    best_off = fr.first;
                    best_size = fr.second;
                }
        if (!best) {
            const size_t want = std::max(bytes, (size_t)2 << 30);
            char *base = nullptr;
            hipError_t e = hipMalloc((void **)&base, want);
            size_t got = want;
            if (e != hipSuccess && want > bytes) {
                (void)hipGetLastError();
                e = hipMalloc((void **)&base, bytes);
                got = bytes;
            }
            if (e != hipSuccess) return e;
            Slab &sl = slabs_[base];
            sl.base = base;
            sl.size = got;
            sl.free[0] = got;
            reserved_ += (int64_t)got;
            best = &sl;
            best_off = 0;
            best_size = got;
        }
        best->free.erase(best_off);
        if (best_size > bytes) best->free[best_off + bytes] = best_size - bytes;
        char *q = best->base + best_off;
        live_[q] = {best, bytes};
        in_use_ += (int64_t)bytes;
        *p = q;
        return hipSuccess;
    }
    void release(void *p) {
        auto it = live_.find((char *)p);
        if (it == live_.end()) return;
        Slab *sl = it->second.first;
        size_t off = (size_t)((char *)p - sl->base), size = it->second.second;
        in_use_ -= (int64_t)size;
        live_.erase(it);
        auto nx = sl->free.lower_bound(off);
        if (nx != sl->free.end() && nx->first == off + size) {   // merge with the next free range
            size += nx->second;
            nx = sl->free.erase(nx);
        }
        if (nx != sl->free.begin()) {                               // and with the previous one
            auto pv = std::prev(nx);
            if (pv->first + pv->second == off) {
                pv->second += size;
                return;
            }
        }
        sl->free[off] = size;
    }
    int64_t in_use() const { return in_use_; }
    int64_t reserved() const { return reserved_; }

  private:
    struct Slab {
        char *base = nullptr;
        size_t size = 0;
        std::map<size_t, size_t> free;   // offset -> size
    };
    std::map<char *, Slab> slabs_;
    std::map<char *, std::pair<Slab *, size_t>> live_;
    int64_t in_use_ = 0, reserved_ = 0;
};

struct Job;
// a job's view of the pool: its grid allocates through this, which also keeps
// the job's own byte count
struct JobMem {
    DevPool *pool = nullptr;
    int64_t bytes = 0;
    std::map<void *, size_t> sizes;
};

hipError_t job_alloc(void *ctx, void **p, size_t bytes) {
    JobMem *m = (JobMem *)ctx;
    const hipError_t e = m->pool->alloc(p, bytes);
    if (e == hipSuccess) {
        const size_t r = (std::max<size_t>(bytes, 1) + 255) / 256 * 256;
        m->sizes[*p] = r;
        m->bytes += (int64_t)r;
    }
    return e;
}

void job_release(void *ctx, void *p) {
    JobMem *m = (JobMem *)ctx;
    auto it = m->sizes.find(p);
    if (it != m->sizes.end()) {
        m->bytes -= (int64_t)it->second;
        m->sizes.erase(it);
    }
    m->pool->release(p);
}

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
    int blocked_rounds = 0;       // rounds passed over while it did not fit
    JobMem mem;
    GridMem gm{};
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
    int64_t budget = 0;
    int quantum = 1, max_active = 0, aging = 8;
    DevPool pool;
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

static void release_job_grid(scp_server *, Job *j) {
    if (j->g) {
        scp_grid_destroy(j->g);   // also leaves its batch; its memory returns to the pool
        j->g = nullptr;
    }
}

// device bytes a job needs at once while it is admitted: the whole grid (emit
// buffers included) plus the projection's five N^3 work arrays, and a margin
// for the pool's 256-byte rounding of each buffer
static int64_t admit_bytes(const Job *j) {
    const int64_t n3 = (int64_t)j->p.N * j->p.N * j->p.N;
    const bool proj = j->project && j->p.mode == 2 && j->p.g != 0.0;
    return scp_grid_bytes(&j->p) + (proj ? 5 * n3 * 8 : 0) + (32 << 10);
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

extern "C" int scp_server_set_aging(scp_server *s, int rounds) {
    if (!s) return set_err(-1, "null server");
    if (rounds < 0) return set_err(-1, "aging=%d must be >= 0", rounds);
    s->aging = rounds;
    return 0;
}

extern "C" int scp_server_submit(scp_server *s, const scp_job *job, int64_t *id) {
    if (!s || !job || !id) return set_err(-1, "null argument");
    *id = 0;
    if (job->params.mode == 0) return set_err(-1, "jobs run the complex or gauged step");
    if (job->nsteps < 1) return set_err(-1, "nsteps=%d must be >= 1", job->nsteps);
    if (job->nballs < 0 || (job->nballs > 0 && !job->balls)) return set_err(-1, "bad ball list");
    if (job->params.step_kernel != 0) return set_err(-1, "jobs run the LDS-staged march (step_kernel 0)");
    if (job->params.bc_type == 1) return set_err(-1, "jobs run the periodic or sponge boundary (bc_type 0 or 2)");
    const int64_t gbytes = scp_grid_bytes(&job->params);
    if (gbytes < 0) return -1;   // message set by the parameter check
    const int64_t n3 = (int64_t)job->params.N * job->params.N * job->params.N;
    const int64_t bytes = gbytes + (job->gauss_project ? 5 * n3 * 8 : 0) + (32 << 10);
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
    j->mem.pool = &s->pool;
    j->gm = GridMem{&j->mem, job_alloc, job_release};
    j->t_submit = s->now();
    s->queue.push_back(j);
    s->jobs[j->id] = j;
    *id = j->id;
    return 0;
}

// create, seed, project and enrol one job; on failure the job is marked failed.
// The emit buffers hold the seed for the initial force + kick and are released
// once the step state is formed.
static int admit(scp_server *s, Job *j) {
    int rc = grid_create_with(&j->p, &j->gm, true, &j->g);
    if (!rc) rc = scp_grid_seed_balls(j->g, j->balls.data(), (int)j->balls.size(), j->theta_sigma, j->link_sigma, j->seed);
    if (!rc && j->project && j->p.mode == 2 && j->p.g != 0.0) {
        int it = 0;
        double gm = 0;
        rc = scp_grid_gauss_project(j->g, &it, &gm);
    }
    if (!rc) rc = grid_emit_buffers(j->g, false);
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
    for (auto it = s->queue.begin(); it != s->queue.end();) {
        if (s->max_active && active >= s->max_active) break;
        Job *j = *it;
        if (s->pool.in_use() + admit_bytes(j) > s->budget) {
            if (j->blocked_rounds >= s->aging) break;   // aged: later jobs wait behind it
            j->blocked_rounds++;
            ++it;
            continue;
        }
        it = s->queue.erase(it);
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
            const int64_t left = j->nsteps - j->done_steps;
            counts[m] = (int)std::min<int64_t>(s->quantum, left);
            if (counts[m] == left) {
                // the job's last chain emits its state: its emit buffers first, if they fit
                const int64_t eb = grid_emit_bytes(&j->p) + (16 << 10);
                if (s->pool.in_use() + eb > s->budget || grid_emit_buffers(j->g, true) != 0)
                    counts[m] = (int)(left - 1);     // one step short: it finishes once memory frees up
            }
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
                if (int e = grid_release_step_state(j->g)) return e;   // the result is in the emit buffers
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
    info->bytes = j->mem.bytes;
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
    st->bytes_in_use = s->pool.in_use();
    st->site_updates = s->site_updates;
    st->busy_seconds = s->busy;
    st->now = s->now();
    return 0;
}
