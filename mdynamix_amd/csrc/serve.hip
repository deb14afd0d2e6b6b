// Serving loop: MD-segment jobs of different sizes admitted by an HBM
// budget and advanced together by fused batched launches.
//
// A worker thread owns every HIP call.  Per scheduling round it
//   1. admits queued jobs in arrival order while a slot is free and the
//      estimated device memory (mdx_estimate_bytes) fits the budget:
//      creates the engine, uploads coordinates and starts the segment
//      (list + first force call), so the job's coordinates and lists stay
//      resident in HBM until it completes;
//   2. re-forms the batch when the set of active jobs changed (a new
//      block map; the force call is captured again as a hipGraph);
//   3. advances every active job by min(quantum, shortest remaining
//      segment) steps with one launch per kernel and step for all jobs;
//   4. retires finished jobs: final state to the caller's arrays, engine
//      destroyed, memory returned to the budget, result queued for poll.
// Jobs submitted while a round runs are admitted at the next round, so new
// work joins the device as soon as slots and memory free up.  Every job's
// arithmetic is independent of which jobs share its batch (per-system
// argument structs, fixed-order per-system reductions), so its result is
// the same alone and in the mix.

#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>

namespace {

struct ServeJob {
    mdx_job_desc d{};
    int64_t id = 0;
    int64_t est = 0;
    double t_submit = 0.0, t_admit = 0.0;
    mdx_engine *h = nullptr;
    int32_t done_steps = 0;
    double pairs = 0.0;
    double last[kTrace] = {};
};

}  // namespace

struct mdx_server {
    int64_t budget = 0;
    int32_t max_slots = 1, quantum = 1;
    std::chrono::steady_clock::time_point t0;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<ServeJob> queue;
    std::vector<mdx_job_result> done;
    std::vector<ServeJob> active;            // worker thread only
    int64_t used = 0;                        // estimated bytes of active jobs
    int64_t next_id = 1;
    int32_t n_active = 0;                    // mirror of active.size() under mu
    bool stop = false;
    std::thread worker;
    std::string err;
    // statistics
    int64_t rounds = 0;
    double busy_s = 0.0;

    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
};

namespace {

void serve_finish(mdx_server *s, ServeJob &j, int32_t status, const std::string &msg) {
    mdx_job_result r{};
    r.id = j.id;
    r.status = status;
    r.nsteps = j.done_steps;
    r.t_submit = j.t_submit;
    r.t_admit = j.t_admit;
    r.t_done = s->now();
    r.epot = j.last[0];
    r.ekin = j.last[1];
    r.temperature = j.last[2];
    r.pair_interactions = (int64_t)std::llround(j.pairs);
    r.bytes = j.est;
    std::snprintf(r.message, sizeof(r.message), "%s", msg.c_str());
    if (j.h) {
        if (status == MDX_OK)
            mdx_md_get_state(j.h, j.d.x_out, j.d.y_out, j.d.z_out, j.d.vx_out, j.d.vy_out, j.d.vz_out);
        mdx_destroy(j.h);
        j.h = nullptr;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    s->done.push_back(r);
}

mdx_status serve_admit(mdx_server *s, ServeJob &j) {
    mdx_engine *h = nullptr;
    mdx_status st = mdx_create(j.d.topology, j.d.params, &h);
    if (st == MDX_OK) st = mdx_set_positions(h, j.d.x, j.d.y, j.d.z);
    if (st == MDX_OK) st = mdx_set_timing(h, 0);
    if (st == MDX_OK) st = mdx_md_init(h, j.d.md, j.d.vx, j.d.vy, j.d.vz);
    if (st != MDX_OK) {
        const std::string msg = h ? h->err : std::string("engine creation failed");
        j.h = h;
        serve_finish(s, j, st, msg);
        return st;
    }
    j.h = h;
    j.t_admit = s->now();
    return MDX_OK;
}

void serve_loop(mdx_server *s) {
    mdx_batch *batch = nullptr;
    std::vector<mdx_engine *> members;
    std::vector<double> trace;
    for (;;) {
        // 1. admission in arrival order
        std::vector<ServeJob> admit;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            if (s->active.empty())
                s->cv.wait(lk, [&] { return s->stop || !s->queue.empty(); });
            if (s->stop) break;
            while (!s->queue.empty() && (int)(s->active.size() + admit.size()) < s->max_slots &&
                   (s->used + s->queue.front().est <= s->budget || (s->active.empty() && admit.empty()))) {
                s->used += s->queue.front().est;
                admit.push_back(std::move(s->queue.front()));
                s->queue.pop_front();
            }
            s->n_active = (int32_t)(s->active.size() + admit.size());
        }
        for (ServeJob &j : admit) {
            if (serve_admit(s, j) == MDX_OK) {
                s->active.push_back(std::move(j));
            } else {
                std::lock_guard<std::mutex> lk(s->mu);
                s->used -= j.est;
            }
        }
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->n_active = (int32_t)s->active.size();
        }
        s->cv.notify_all();
        if (s->active.empty()) continue;
        // 2. batch of the active jobs
        std::vector<mdx_engine *> now;
        for (ServeJob &j : s->active) now.push_back(j.h);
        if (now != members) {
            if (batch) mdx_batch_destroy(batch);
            batch = nullptr;
            members = now;
            if (mdx_batch_create(members.data(), (int32_t)members.size(), &batch) != MDX_OK) {
                for (ServeJob &j : s->active) serve_finish(s, j, MDX_ERR_HIP, "batch creation failed");
                s->active.clear();
                members.clear();
                if (batch) mdx_batch_destroy(batch);
                batch = nullptr;
                continue;
            }
        }
        // 3. advance all active jobs together
        int q = s->quantum;
        for (ServeJob &j : s->active) q = std::min(q, j.d.nsteps - j.done_steps);
        const size_t K = s->active.size();
        trace.assign(K * (size_t)std::max(q, 1) * kTrace, 0.0);
        const double tb = s->now();
        const mdx_status st = q > 0 ? mdx_batch_md_run(batch, q, trace.data()) : MDX_OK;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->busy_s += s->now() - tb;
            s->rounds += 1;
        }
        if (st != MDX_OK) {
            const std::string msg = mdx_batch_last_error(batch);
            for (ServeJob &j : s->active) serve_finish(s, j, st, msg);
            s->active.clear();
        } else {
            for (size_t k = 0; k < K; ++k) {
                ServeJob &j = s->active[k];
                const double *t = trace.data() + k * (size_t)q * kTrace;
                for (int i = 0; i < q; ++i) j.pairs += t[(size_t)i * kTrace + 4];
                if (q > 0)
                    for (int c = 0; c < kTrace; ++c) j.last[c] = t[(size_t)(q - 1) * kTrace + c];
                j.done_steps += q;
            }
        }
        // 4. retire finished jobs
        std::vector<ServeJob> keep;
        int64_t freed = 0;
        for (ServeJob &j : s->active) {
            if (j.done_steps >= j.d.nsteps) {
                freed += j.est;
                serve_finish(s, j, MDX_OK, "");
            } else {
                keep.push_back(std::move(j));
            }
        }
        s->active.swap(keep);
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->used -= freed;
            s->n_active = (int32_t)s->active.size();
        }
        s->cv.notify_all();
    }
    if (batch) mdx_batch_destroy(batch);
    for (ServeJob &j : s->active) serve_finish(s, j, MDX_ERR_STATE, "server stopped");
    s->active.clear();
}

}  // namespace

extern "C" {

mdx_status mdx_serve_create(int64_t hbm_budget_bytes, int32_t max_slots, int32_t quantum, mdx_server **out) {
    if (!out || hbm_budget_bytes <= 0 || max_slots <= 0 || quantum <= 0) return MDX_ERR_ARG;
    mdx_server *s = new mdx_server();
    s->budget = hbm_budget_bytes;
    s->max_slots = max_slots;
    s->quantum = quantum;
    s->t0 = std::chrono::steady_clock::now();
    s->worker = std::thread(serve_loop, s);
    *out = s;
    return MDX_OK;
}

mdx_status mdx_serve_submit(mdx_server *s, const mdx_job_desc *d, int64_t *id) {
    if (!s || !d || !d->topology || !d->params || !d->md || !d->x || !d->y || !d->z || !d->vx || !d->vy ||
        !d->vz || d->nsteps < 0)
        return MDX_ERR_ARG;
    ServeJob j;
    j.d = *d;
    const mdx_status st = mdx_estimate_bytes(d->topology, d->params, 1, &j.est);
    if (st != MDX_OK) return st;
    if (j.est > s->budget) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->err = "job needs more device memory than the whole budget";
        return MDX_ERR_NOMEM;
    }
    {
        std::lock_guard<std::mutex> lk(s->mu);
        j.id = s->next_id++;
        j.t_submit = s->now();
        if (id) *id = j.id;
        s->queue.push_back(std::move(j));
    }
    s->cv.notify_all();
    return MDX_OK;
}

mdx_status mdx_serve_poll(mdx_server *s, mdx_job_result *out, int32_t max, int32_t *n, int32_t *active,
                          int32_t *queued) {
    if (!s || (max > 0 && !out)) return MDX_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    const int32_t k = std::min<int32_t>(max, (int32_t)s->done.size());
    for (int32_t i = 0; i < k; ++i) out[i] = s->done[i];
    s->done.erase(s->done.begin(), s->done.begin() + k);
    if (n) *n = k;
    if (active) *active = s->n_active;
    if (queued) *queued = (int32_t)s->queue.size();
    return MDX_OK;
}

mdx_status mdx_serve_wait(mdx_server *s, double timeout_s) {
    if (!s) return MDX_ERR_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    const bool ok = s->cv.wait_for(lk, std::chrono::duration<double>(timeout_s),
                                   [&] { return s->queue.empty() && s->n_active == 0; });
    return ok ? MDX_OK : MDX_ERR_STATE;
}

mdx_status mdx_serve_stats(mdx_server *s, int64_t *rounds, double *busy_s, double *clock_s) {
    if (!s) return MDX_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    if (rounds) *rounds = s->rounds;
    if (busy_s) *busy_s = s->busy_s;
    if (clock_s) *clock_s = s->now();
    return MDX_OK;
}

const char *mdx_serve_last_error(const mdx_server *s) { return s ? s->err.c_str() : "null server"; }

mdx_status mdx_serve_destroy(mdx_server *s) {
    if (!s) return MDX_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv.notify_all();
    if (s->worker.joinable()) s->worker.join();
    delete s;
    return MDX_OK;
}

}  // extern "C"
