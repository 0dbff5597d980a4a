// server.hip -- serving loop: solves ("jobs") of different sizes arrive over
// time, are admitted while their device memory fits a budget, stay resident
// on the device until they finish, and advance together.
//
// Continuous batching: the loop works in rounds.  Each round admits the jobs
// that have arrived, in arrival order, while the head of the queue fits the
// budget (and the active-job limit); then every active job advances by up
// to `quantum` steps -- jobs that qualify for the fused step all in one
// batch (one row and one column launch per size class per step), the others
// each on its own stream with its general step program replayed from a
// hipGraph -- so the work of all active jobs overlaps on the device.  Jobs
// that reach their step count are handed back (done callback), destroyed,
// and their memory is released for the next admission.  A job's result
// does not depend on the other jobs: systems share no state and every
// system keeps its own step counter and noise stream.
//
// No counterpart in the reference, which advances one evolver at a time
// (src/evolver.cpp:199-226).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <deque>
#include <thread>
#include <vector>

#include "system_impl.hpp"

using namespace cupss;

namespace {

struct Job {
    double arrival = 0, admitted = -1, finished = -1;
    int nsteps = 0, left = 0;
    size_t estimate = 0, bytes = 0;
    cupss_job_build build = nullptr;
    cupss_job_done done = nullptr;
    void *user = nullptr;
    cupss_system sys = nullptr;
    bool batched = false;
    double points = 0;
};

}  // namespace

struct cupss_server_t {
    size_t budget = 0;
    int quantum = 1;
    int max_active = 1;
    int graph_steps = 0;
    std::vector<Job> jobs;
    cupss_batch batch = nullptr;
    double wall = 0, point_steps = 0;
    size_t peak_bytes = 0;
    int peak_active = 0;
    long rounds = 0;
};

extern "C" int cupss_server_create(size_t budget_bytes, int quantum, int max_active, cupss_server *out)
{
    if (!out || quantum < 1 || max_active < 1) return fail(CUPSS_ERR_INVALID, "bad server parameters");
    cupss_server v = new cupss_server_t();
    v->budget = budget_bytes;
    v->quantum = quantum;
    v->max_active = max_active;
    v->graph_steps = quantum >= 2 ? (quantum & ~1) : 0;
    if (int r = cupss_batch_create(nullptr, &v->batch)) {
        delete v;
        return r;
    }
    cupss_batch_set_autotune(v->batch, 0);  // membership changes every few rounds
    if (v->graph_steps) cupss_batch_set_graph_steps(v->batch, v->graph_steps);
    *out = v;
    return CUPSS_OK;
}

extern "C" int cupss_server_destroy(cupss_server v)
{
    if (!v) return CUPSS_OK;
    if (v->batch) cupss_batch_destroy(v->batch);
    for (auto &j : v->jobs)
        if (j.sys) cupss_system_destroy(j.sys);
    delete v;
    return CUPSS_OK;
}

extern "C" int cupss_server_submit(cupss_server v, double arrival, int nsteps, size_t bytes_estimate,
                                   cupss_job_build build, cupss_job_done done, void *user, int *job)
{
    if (!v || !build || nsteps < 1) return fail(CUPSS_ERR_INVALID, "bad job");
    Job j;
    j.arrival = arrival;
    j.nsteps = j.left = nsteps;
    j.estimate = bytes_estimate;
    j.build = build;
    j.done = done;
    j.user = user;
    v->jobs.push_back(j);
    if (job) *job = (int)v->jobs.size() - 1;
    return CUPSS_OK;
}

static int admit(cupss_server v, int id, double now)
{
    Job &j = v->jobs[id];
    cupss_system s = nullptr;
    if (int r = j.build(j.user, id, &s)) return fail(r, "job build callback failed");
    if (!s) return fail(CUPSS_ERR_INVALID, "job build callback returned no system");
    j.sys = s;
    if (int r = cupss_system_device_bytes(s, &j.bytes)) return r;
    j.points = (double)s->g.n;
    j.admitted = now;
    if (cupss_system_fused_eligible(s)) {
        if (int r = cupss_batch_add(v->batch, s)) return r;
        j.batched = true;
    } else {
        if (int r = cupss_system_set_graph_steps(s, v->graph_steps)) return r;
    }
    return CUPSS_OK;
}

static int retire(cupss_server v, int id, double now)
{
    Job &j = v->jobs[id];
    if (j.batched)
        if (int r = cupss_batch_remove(v->batch, j.sys)) return r;
    HIP_TRY(hipStreamSynchronize(j.sys->stream));
    j.finished = now;
    if (j.done) j.done(j.user, id, j.sys);
    cupss_system_destroy(j.sys);
    j.sys = nullptr;
    return CUPSS_OK;
}

extern "C" int cupss_server_run(cupss_server v)
{
    if (!v) return fail(CUPSS_ERR_INVALID, "null server");
    using clock = std::chrono::steady_clock;
    const auto t0 = clock::now();
    auto now = [&]() { return std::chrono::duration<double>(clock::now() - t0).count(); };
    std::vector<int> order;
    for (int i = 0; i < (int)v->jobs.size(); i++)
        if (v->jobs[i].finished < 0) order.push_back(i);
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return v->jobs[a].arrival < v->jobs[b].arrival; });
    size_t next = 0, used = 0;
    std::deque<int> queue;
    std::vector<int> active;
    size_t left = order.size();
    while (left > 0) {
        const double t = now();
        while (next < order.size() && v->jobs[order[next]].arrival <= t) queue.push_back(order[next++]);
        // admission: arrival order; an idle device takes the head even if it
        // alone exceeds the budget
        while (!queue.empty() && (int)active.size() < v->max_active) {
            const Job &h = v->jobs[queue.front()];
            if (!active.empty() && used + h.estimate > v->budget) break;
            const int id = queue.front();
            queue.pop_front();
            if (int r = admit(v, id, now())) return r;
            used += v->jobs[id].bytes;
            active.push_back(id);
        }
        v->peak_bytes = std::max(v->peak_bytes, used);
        v->peak_active = std::max(v->peak_active, (int)active.size());
        if (active.empty()) {
            if (next < order.size()) {
                const double wait = v->jobs[order[next]].arrival - now();
                if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 0.001)));
            }
            continue;
        }
        // one round: the batch advances by the smallest remaining count of
        // its members (at most the quantum), the others by their own
        int qb = v->quantum;
        bool any_batched = false;
        for (int id : active)
            if (v->jobs[id].batched) {
                qb = std::min(qb, v->jobs[id].left);
                any_batched = true;
            }
        if (any_batched) {
            if (int r = cupss_batch_advance(v->batch, qb)) return r;
        }
        for (int id : active) {
            Job &j = v->jobs[id];
            if (j.batched) continue;
            const int q = std::min(v->quantum, j.left);
            if (int r = cupss_system_advance(j.sys, q)) return r;
        }
        for (int id : active) {
            Job &j = v->jobs[id];
            const int q = j.batched ? qb : std::min(v->quantum, j.left);
            j.left -= q;
            v->point_steps += j.points * q;
        }
        if (any_batched)
            if (int r = cupss_batch_synchronize(v->batch)) return r;
        for (int id : active)
            if (!v->jobs[id].batched) HIP_TRY(hipStreamSynchronize(v->jobs[id].sys->stream));
        v->rounds++;
        std::vector<int> still;
        for (int id : active) {
            if (v->jobs[id].left > 0) {
                still.push_back(id);
                continue;
            }
            used -= std::min(used, v->jobs[id].bytes);
            if (int r = retire(v, id, now())) return r;
            left--;
        }
        active.swap(still);
    }
    v->wall = now();
    return CUPSS_OK;
}

extern "C" int cupss_server_job_times(cupss_server v, int job, double *arrival, double *admitted, double *finished)
{
    if (!v || job < 0 || job >= (int)v->jobs.size()) return fail(CUPSS_ERR_INVALID, "bad job id");
    const Job &j = v->jobs[job];
    if (arrival) *arrival = j.arrival;
    if (admitted) *admitted = j.admitted;
    if (finished) *finished = j.finished;
    return CUPSS_OK;
}

extern "C" int cupss_server_stats(cupss_server v, double *wall_s, double *point_steps, size_t *peak_bytes,
                                  int *peak_active)
{
    if (!v) return fail(CUPSS_ERR_INVALID, "null server");
    if (wall_s) *wall_s = v->wall;
    if (point_steps) *point_steps = v->point_steps;
    if (peak_bytes) *peak_bytes = v->peak_bytes;
    if (peak_active) *peak_active = v->peak_active;
    return CUPSS_OK;
}
