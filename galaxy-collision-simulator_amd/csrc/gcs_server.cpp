// gcs_server.cpp -- served form of the step: continuous batching of collision
// jobs on one engine (host C++ over the gcs_* C ABI).
//
// Jobs arrive on a schedule; worker threads generate each job's initial
// conditions when it arrives (or, with no worker threads, the engine
// generates them on the device when the job is admitted); the stepping loop
// admits generated jobs in arrival order while their particles fit the HBM
// budget (and the job cap),
// advances every resident job together -- one launch sequence per step over
// the whole forest -- in rounds that end when the next job finishes or after
// steps_per_round steps, and retires finished jobs.  Each job's result does
// not depend on what else is resident: every instance has its own bounds,
// sort range, tree and traversal groups.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "gcs.h"
#include "gcs_common.hpp"
#include "gcs_error.hpp"

namespace {

struct Job {
    gcs_job spec;
    std::vector<gcs_star> stars;
    bool queued = false;
    bool ready = false;
    bool admitted = false;
    bool finished = false;
    int32_t remaining = 0;
    int32_t iid = -1;
    double admit_s = 0.0, finish_s = 0.0;
};

double now_s(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

struct gcs_server {
    gcs_server_config cfg;
    gcs_engine* eng = nullptr;
    int64_t cap_particles = 0;
    int32_t cap_jobs = 0;

    std::vector<Job> jobs;  // kept in arrival order
    std::mutex mu;
    std::condition_variable cv_gen, cv_ready;
    std::deque<size_t> gen_queue;
    std::vector<std::thread> workers;
    bool stop = false;
    int gen_error = GCS_OK;

    std::vector<size_t> resident;
    int64_t resident_particles = 0;
    // jobs[0, next_release) have arrived; jobs[0, next_gen) are queued for
    // generation (a bounded look-ahead past the admission head keeps host
    // memory bounded under overload); jobs[0, next_admit) are admitted.
    size_t next_release = 0, next_gen = 0, next_admit = 0, lookahead = 32;
    bool device_gen = false;  // no generator threads: jobs are generated on the device
    std::vector<gcs_job_result> results;
    std::map<int64_t, std::vector<gcs_star>> finals;
    double particle_steps = 0.0, run_s = 0.0, resident_sum = 0.0, gpu_s = 0.0;
    double full_t0 = -1.0, full_steps = 0.0, full_s = 0.0;  // from the first full round
    bool full = false;  // admission was held back by the cap in this run
    // The server clock: arrival, admission and finish times count from the
    // start of the first gcs_server_run; later runs continue it.
    std::chrono::steady_clock::time_point clock0;
    bool clock_started = false;
    int64_t steps_run = 0, rounds = 0, admitted = 0;

    void worker()
    {
        for (;;) {
            size_t j;
            gcs_job js;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_gen.wait(lk, [&] { return stop || !gen_queue.empty(); });
                if (stop) return;
                j = gen_queue.front();
                gen_queue.pop_front();
                js = jobs[j].spec;
            }
            gcs_collision_spec spec;
            gcs_spec_random(js.seed, js.id, js.n, js.n, &spec);
            std::vector<gcs_star> st;
            int rc = GCS_OK;
            try {
                st.resize((size_t)js.n);
                rc = gcs_generate(&spec, st.data());
            } catch (...) {
                rc = GCS_ERR_OOM;
            }
            {
                std::lock_guard<std::mutex> lk(mu);
                if (rc != GCS_OK) gen_error = rc;
                jobs[j].stars.swap(st);
                jobs[j].ready = true;
            }
            cv_ready.notify_all();
        }
    }
};

namespace {

int admit(gcs_server* s, double t)
{
    while (s->next_admit < s->next_release) {
        Job* j = nullptr;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            if (s->gen_error != GCS_OK) return s->gen_error;
            if (!s->jobs[s->next_admit].ready) break;
            j = &s->jobs[s->next_admit];
        }
        if ((int32_t)s->resident.size() >= s->cap_jobs ||
            s->resident_particles + j->spec.n > s->cap_particles) {
            s->full = true;
            break;
        }
        int32_t iid = -1;
        int rc;
        if (s->device_gen) {
            gcs_collision_spec spec;
            gcs_spec_random(j->spec.seed, j->spec.id, j->spec.n, j->spec.n, &spec);
            rc = gcs_instance_add_generated(s->eng, &spec, &iid);
        } else {
            rc = gcs_instance_add(s->eng, j->stars.data(), j->spec.n, &iid);
        }
        if (rc != GCS_OK) return rc;
        std::vector<gcs_star>().swap(j->stars);
        j->iid = iid;
        j->admitted = true;
        j->admit_s = t;
        j->remaining = j->spec.steps;
        s->resident.push_back(s->next_admit);
        s->resident_particles += j->spec.n;
        ++s->admitted;
        ++s->next_admit;
    }
    return GCS_OK;
}

}  // namespace

extern "C" {

int gcs_server_create(const gcs_server_config* c, gcs_server** out)
{
    using gcs_detail::set_error;
    if (!c || !out) return set_error(GCS_ERR_INVALID, "null argument");
    *out = nullptr;
    if (c->max_jobs < 1 || c->steps_per_round < 1 || c->hbm_budget_bytes <= 0)
        return set_error(GCS_ERR_INVALID, "server config: max_jobs, steps_per_round and "
                                          "hbm_budget_bytes must be positive");
    gcs_server* s = new (std::nothrow) gcs_server();
    if (!s) return GCS_ERR_OOM;
    s->cfg = *c;
    // Particle capacity from the budget: the engine's bytes grow linearly in
    // the particle count (plus a fixed part for the traversal wave slots).
    const int32_t kcap = std::max<int32_t>(1, std::min<int32_t>(c->max_jobs, 1 << 16));
    const int64_t b1 = gcs_device_bytes(1 << 20, kcap), b2 = gcs_device_bytes(1 << 21, kcap);
    const int64_t per = std::max<int64_t>(1, (b2 - b1) >> 20);
    const int64_t fixed = std::max<int64_t>(0, b1 - per * (1 << 20)) + ((int64_t)1 << 30);
    int64_t cap = (c->hbm_budget_bytes - fixed) / per;
    if (c->max_resident > 0) cap = std::min(cap, c->max_resident);
    cap = std::min<int64_t>(cap, gcs::kMaxParticles);
    if (cap < 1) {
        delete s;
        return set_error(GCS_ERR_CAPACITY, "HBM budget holds no particles");
    }
    s->cap_particles = cap;
    s->cap_jobs = kcap;
    int rc = gcs_create(cap, kcap, 0, &s->eng);
    if (rc != GCS_OK) {
        delete s;
        return rc;
    }
    s->device_gen = c->gen_threads <= 0;
    const int nt = s->device_gen ? 0 : std::min(32, c->gen_threads);
    s->lookahead = (size_t)std::max(32, 4 * nt);
    for (int t = 0; t < nt; ++t) s->workers.emplace_back([s] { s->worker(); });
    *out = s;
    return GCS_OK;
}

int gcs_server_destroy(gcs_server* s)
{
    if (!s) return GCS_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_gen.notify_all();
    for (auto& t : s->workers) t.join();
    gcs_destroy(s->eng);
    delete s;
    return GCS_OK;
}

int gcs_server_submit(gcs_server* s, const gcs_job* jobs, int32_t n)
{
    using gcs_detail::set_error;
    if (!s || (!jobs && n > 0) || n < 0) return set_error(GCS_ERR_INVALID, "bad job array");
    std::lock_guard<std::mutex> lk(s->mu);
    for (int32_t i = 0; i < n; ++i)
        if (jobs[i].n < 2 || jobs[i].steps < 1 || jobs[i].n > s->cap_particles)
            return set_error(jobs[i].n > s->cap_particles ? GCS_ERR_CAPACITY : GCS_ERR_INVALID,
                             "job needs n >= 2, steps >= 1 and n within the resident cap");
    for (int32_t i = 0; i < n; ++i) {
        Job j;
        j.spec = jobs[i];
        s->jobs.push_back(std::move(j));
    }
    // Arrival order among the jobs not yet released.
    std::stable_sort(s->jobs.begin() + (long)s->next_release, s->jobs.end(),
                     [](const Job& a, const Job& b) { return a.spec.arrival_s < b.spec.arrival_s; });
    return GCS_OK;
}

int gcs_server_run(gcs_server* s, double duration_s)
{
    if (!s) return GCS_ERR_INVALID;
    if (!s->clock_started) {
        s->clock0 = std::chrono::steady_clock::now();
        s->clock_started = true;
    }
    const auto t0 = s->clock0;
    const double t_begin = now_s(t0);
    int rc = GCS_OK;
    for (;;) {
        const double t = now_s(t0);
        // Release arrivals; hand the look-ahead window to the generators.
        {
            std::lock_guard<std::mutex> lk(s->mu);
            while (s->next_release < s->jobs.size() &&
                   s->jobs[s->next_release].spec.arrival_s <= t)
                ++s->next_release;
            while (s->next_gen < s->next_release && s->next_gen < s->next_admit + s->lookahead) {
                s->jobs[s->next_gen].queued = true;
                if (s->device_gen) s->jobs[s->next_gen].ready = true;  // generated at admission
                else s->gen_queue.push_back(s->next_gen);
                ++s->next_gen;
            }
        }
        s->cv_gen.notify_all();
        rc = admit(s, t);
        if (rc != GCS_OK) break;
        const bool all_done = s->next_admit == s->jobs.size() && s->resident.empty();
        if (all_done && duration_s <= 0) break;
        if (duration_s > 0 && t - t_begin >= duration_s) break;
        if (s->resident.empty()) {
            // Wait for the next arrival or generation.
            std::unique_lock<std::mutex> lk(s->mu);
            double wait = 0.005;
            if (s->next_release < s->jobs.size())
                wait = std::min(wait, std::max(0.0, s->jobs[s->next_release].spec.arrival_s - t));
            s->cv_ready.wait_for(lk, std::chrono::duration<double>(wait));
            continue;
        }
        int32_t k = s->cfg.steps_per_round;
        for (size_t r : s->resident) k = std::min(k, s->jobs[r].remaining);
        const double ts = now_s(t0);
        if (s->full && s->full_t0 < 0.0) s->full_t0 = ts;
        rc = gcs_step(s->eng, &s->cfg.params, k);
        if (rc == GCS_OK) rc = gcs_sync(s->eng);
        if (rc != GCS_OK) break;
        s->gpu_s += now_s(t0) - ts;
        if (s->full_t0 >= 0.0) s->full_steps += (double)s->resident_particles * k;
        s->particle_steps += (double)s->resident_particles * k;
        s->resident_sum += (double)s->resident_particles * k;
        s->steps_run += k;
        ++s->rounds;
        const double tf = now_s(t0);
        std::vector<size_t> keep;
        for (size_t r : s->resident) {
            Job& j = s->jobs[r];
            j.remaining -= k;
            if (j.remaining > 0) {
                keep.push_back(r);
                continue;
            }
            if (s->cfg.keep_final) {
                std::vector<gcs_star> st((size_t)j.spec.n);
                rc = gcs_read_instance(s->eng, j.iid, st.data());
                if (rc != GCS_OK) break;
                s->finals[j.spec.id].swap(st);
            }
            rc = gcs_instance_remove(s->eng, j.iid);
            if (rc != GCS_OK) break;
            j.finished = true;
            j.finish_s = tf;
            s->resident_particles -= j.spec.n;
            gcs_job_result res;
            res.id = j.spec.id;
            res.n = j.spec.n;
            res.steps = j.spec.steps;
            res.arrival_s = j.spec.arrival_s;
            res.admit_s = j.admit_s;
            res.finish_s = tf;
            s->results.push_back(res);
        }
        if (rc != GCS_OK) break;
        s->resident.swap(keep);
    }
    const double t_end = now_s(t0);
    s->run_s += t_end - t_begin;
    if (s->full_t0 >= 0.0) s->full_s += t_end - s->full_t0;
    s->full_t0 = -1.0;
    s->full = false;
    return rc;
}

int gcs_server_results(gcs_server* s, gcs_job_result* out, int32_t cap, int32_t* n_out)
{
    if (!s || !n_out) return GCS_ERR_INVALID;
    const int32_t n = (int32_t)std::min<size_t>(s->results.size(), (size_t)std::max(cap, 0));
    if (out && n > 0) std::memcpy(out, s->results.data(), sizeof(gcs_job_result) * (size_t)n);
    *n_out = (int32_t)s->results.size();
    return GCS_OK;
}

int gcs_server_stats(gcs_server* s, double* out, int32_t n)
{
    if (!s || (!out && n > 0) || n < 0) return GCS_ERR_INVALID;
    const double v[11] = {s->particle_steps,
                          s->run_s,
                          (double)s->results.size(),
                          (double)s->admitted,
                          (double)(s->next_release - s->next_admit),
                          s->steps_run ? s->resident_sum / (double)s->steps_run : 0.0,
                          (double)s->rounds,
                          (double)s->cap_particles,
                          s->gpu_s,
                          s->full_steps,
                          s->full_s};
    for (int32_t i = 0; i < n && i < 11; ++i) out[i] = v[i];
    for (int32_t i = 11; i < n; ++i) out[i] = 0.0;
    return GCS_OK;
}

int gcs_server_final(gcs_server* s, int64_t id, gcs_star* out, int32_t n)
{
    if (!s || !out) return GCS_ERR_INVALID;
    auto it = s->finals.find(id);
    if (it == s->finals.end()) return GCS_ERR_NOT_FOUND;
    if ((size_t)n != it->second.size()) return GCS_ERR_INVALID;
    std::memcpy(out, it->second.data(), sizeof(gcs_star) * (size_t)n);
    return GCS_OK;
}

}  // extern "C"
