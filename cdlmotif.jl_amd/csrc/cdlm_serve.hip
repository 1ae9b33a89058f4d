// cdlm_serve.hip -- continuous-batching server of CDL training jobs (host code over the C ABI).
//
// The reference trains datasets one at a time from the host (find_motifs_fasta ->
// try_to_find_motif -> cdl_optimize, src/WRAPPER/wrap.jl:4-12,140-149).  Here a server keeps
// many jobs resident on one device.  By default they run concurrently, each at its own pace:
// a dispatcher thread admits jobs (as below) and hands each to one of a few worker threads,
// each with its own context and stream, which creates the job's problem in its arena and runs
// its outer iterations back to back; the device interleaves the work-groups of the concurrent
// jobs' kernels, so a small job beside a large one is not held to the large one's iteration
// time, and a job's upload and initialisation overlap the others' iterations.  In lockstep
// mode (opts.lockstep) one server thread advances all resident jobs together instead:
//   * submit copies a job's data to host memory and queues it (arrival time recorded);
//     a job whose device footprint exceeds the whole budget is streamed: admitted when nothing
//     else is resident, it runs alone as a problem whose state lives in pinned host memory and
//     whose sequences move through a device window of the budget in chunks (its ticks are
//     cdlm_optimize of one outer iteration), with the results it gets resident;
//   * between ticks, queued jobs are admitted in arrival order while their footprint
//     (cdlm_footprint_bytes, the exact bytes a problem holds) fits the free budget and the
//     resident jobs hold fewer sequences than fill_seqs (the sequences that occupy every
//     CU: each sequence is one work-group of every per-sequence kernel, so more resident
//     work only shares the device between jobs and delays each of them).  Candidates are
//     taken in arrival order (policy FIFO) or least work first, arrival order breaking ties
//     (policy SHORTEST; work = iter_num x N x Lg x M, proportional to the job's HBM bytes);
//     a candidate that does not fit may be passed by a later one (backfill), but a job
//     passed kMaxBypass times goes first and nothing passes it any more;
//   * a tick runs one outer iteration of the reference schedule (csc_iter CSC iterations,
//     update_best, cdl_iter dictionary iterations; wrap.jl:6-11) for every resident job
//     as one batch (one launch per kernel per shape class, replayed from a HIP graph while
//     the membership is unchanged);
//   * a job that completed its iter_num outer iterations copies its results (objective
//     trace, code threshold, optionally the X_best slice) to host memory; its device memory
//     (one arena per job) goes to a pool that later admissions take from (the smallest arena
//     that holds the job), so device memory is freed only when a new job fits the budget
//     but no pooled arena holds it.  Freeing device memory synchronises the device and
//     leaves driver work that slows the next seconds of kernels, so recycling it keeps
//     both out of the ticks.
// A job's arithmetic does not depend on which other jobs share its ticks, so its results
// are the ones it gets alone.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "cdlm_internal.h"

namespace {

constexpr int kMaxBypass = 4;
using Clock = std::chrono::steady_clock;

struct Job {
    uint64_t id = 0;
    // spec (host copies)
    std::vector<int32_t> fl;
    cdlm_params prm{};
    std::vector<float> s, D0;
    int32_t L4 = 0, N = 0, iter_num = 0, keep_codes = 0;
    uint64_t seed = 0;
    double mult = 1.0;
    size_t footprint = 0;
    double work = 0.0;     // iter_num * N * Lg * M
    bool streamed = false; // footprint beyond the budget: streamed through a device window, alone
    // state
    int32_t state = CDLM_JOB_QUEUED;
    int bypassed = 0;      // later arrivals admitted before this job
    cdlm_problem *p = nullptr;
    void *arena = nullptr; // device memory of the job's problem (recycled or new)
    size_t arena_bytes = 0;
    int outer_done = 0;
    Clock::time_point t_submit, t_admit, t_done;
    // results
    std::vector<double> trace;
    std::vector<float> codes;
    double thresh = 0.0;
    int64_t count = 0;
    float obj_best = 0.f;
    int32_t Lg = 0, gamma = 0;
    std::string error;
};

}  // namespace

struct cdlm_server {
    cdlm_ctx *ctx = nullptr;
    cdlm_server_opts opts{};
    size_t budget = 0;
    long fill = 0;                    // resident sequences beyond which no job is admitted
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::map<uint64_t, std::unique_ptr<Job>> jobs;
    std::deque<uint64_t> queue;
    std::vector<Job *> active;        // server thread only (under mu when read by clients)
    size_t in_use = 0;                // bytes of the resident jobs' arenas
    std::vector<std::pair<void *, size_t>> pool;   // arenas of finished jobs, kept for reuse
    size_t pool_bytes = 0;
    uint64_t next_id = 1;
    bool stop = false;
    std::thread th;
    std::vector<double> latencies;    // completed jobs, seconds
    int64_t ticks = 0;
    std::string last_error;
    // server thread only (lockstep)
    cdlm_batch *batch = nullptr;
    bool membership_changed = true;
    // concurrent mode: workers, jobs handed to them, idle workers, and a counter of events that
    // can make another admission possible (a submission, a job leaving the device)
    std::vector<std::thread> workers;
    std::deque<Job *> ready;
    std::condition_variable cv_ready;
    int idle = 0;
    uint64_t gen = 0;
};

namespace {

void set_error(cdlm_server *srv, const std::string &m) {
    std::lock_guard<std::mutex> lk(srv->mu);
    srv->last_error = m;
}

std::string ctx_error(cdlm_server *srv) {
    const char *e = cdlm_last_error(srv->ctx);
    return e ? std::string(e) : std::string();
}

void fail_job(cdlm_server *srv, Job *j, const std::string &why) {
    std::lock_guard<std::mutex> lk(srv->mu);
    j->state = CDLM_JOB_FAILED;
    j->error = why;
    j->t_done = Clock::now();
    srv->cv_done.notify_all();
}

// Admission (server thread, under mu): returns the jobs to create, each with a recycled arena
// (the smallest pooled one that holds it) or none (a new one of its footprint is allocated);
// pooled arenas that must be freed to make room for new ones are moved to *to_free.  The
// device memory held, resident arenas plus the pool, never exceeds the budget.
std::vector<Job *> pick_admissions(cdlm_server *srv, std::vector<void *> *to_free) {
    std::vector<Job *> out;
    int slots = srv->opts.max_active > 0 ? srv->opts.max_active - (int)srv->active.size() : 1 << 30;
    if (!srv->opts.lockstep) slots = std::min(slots, srv->idle);   // one job per idle worker
    long seqs = 0;
    for (Job *j : srv->active) seqs += j->N;
    // candidate order: jobs passed kMaxBypass times first (arrival order), then by policy
    std::vector<Job *> cand;
    for (uint64_t id : srv->queue) cand.push_back(srv->jobs[id].get());
    auto starving = [](const Job *j) { return j->bypassed >= kMaxBypass; };
    std::stable_sort(cand.begin(), cand.end(), [&](const Job *a, const Job *b) {
        const bool sa = starving(a), sb = starving(b);
        if (sa != sb) return sa;
        if (sa || srv->opts.policy != CDLM_POLICY_SHORTEST) return false;   // keep arrival order
        return a->work < b->work;
    });
    bool blocked = false;   // a starving job does not fit: nothing passes it
    for (Job *j : srv->active)
        if (j->streamed) return out;   // a streamed job has the device to itself
    for (Job *j : cand) {
        if (slots <= 0 || (srv->fill > 0 && seqs >= srv->fill) || blocked) break;
        if (j->streamed) {
            // alone: admitted when nothing is resident, with the whole budget (pooled arenas freed)
            if (out.empty() && srv->active.empty()) {
                for (auto &a : srv->pool) to_free->push_back(a.first);
                srv->pool.clear();
                srv->pool_bytes = 0;
                j->arena = nullptr;
                j->arena_bytes = srv->budget;
                srv->in_use += j->arena_bytes;
                j->state = CDLM_JOB_RUNNING;
                out.push_back(j);
                break;
            }
            if (starving(j)) blocked = true;
            continue;
        }
        int best = -1;
        for (size_t i = 0; i < srv->pool.size(); ++i)
            if (srv->pool[i].second >= j->footprint && (best < 0 || srv->pool[i].second < srv->pool[(size_t)best].second))
                best = (int)i;
        bool admit = false;
        if (best >= 0) {
            j->arena = srv->pool[(size_t)best].first;
            j->arena_bytes = srv->pool[(size_t)best].second;
            srv->pool.erase(srv->pool.begin() + best);
            srv->pool_bytes -= j->arena_bytes;
            admit = true;
        } else if (srv->in_use + j->footprint <= srv->budget) {
            // free pooled arenas, largest first, until a new one of the footprint fits
            while (srv->in_use + srv->pool_bytes + j->footprint > srv->budget && !srv->pool.empty()) {
                auto big = std::max_element(srv->pool.begin(), srv->pool.end(),
                                            [](const std::pair<void *, size_t> &a, const std::pair<void *, size_t> &b) {
                                                return a.second < b.second;
                                            });
                to_free->push_back(big->first);
                srv->pool_bytes -= big->second;
                srv->pool.erase(big);
            }
            j->arena = nullptr;
            j->arena_bytes = j->footprint;
            admit = true;
        }
        if (admit) {
            srv->in_use += j->arena_bytes;
            seqs += j->N;
            --slots;
            j->state = CDLM_JOB_RUNNING;
            out.push_back(j);
        } else if (starving(j)) {
            blocked = true;
        }
    }
    for (Job *j : out) {
        srv->queue.erase(std::find(srv->queue.begin(), srv->queue.end(), j->id));
        for (uint64_t id : srv->queue) {
            Job *w = srv->jobs[id].get();
            if (w->id < j->id) ++w->bypassed;   // an earlier arrival was passed
        }
    }
    return out;
}

// Gives a job's arena back: to the pool (reuse = true) or to the device.
void return_arena(cdlm_server *srv, Job *j, bool reuse) {
    void *a = j->arena;
    const size_t b = j->arena_bytes;
    j->arena = nullptr;
    j->arena_bytes = 0;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->in_use -= b;
        if (a && reuse) {
            srv->pool.emplace_back(a, b);
            srv->pool_bytes += b;
            return;
        }
    }
    if (a) (void)hipFree(a);
}

void finish_job(cdlm_server *srv, Job *j, cdlm_ctx *pctx = nullptr) {
    double th = 0.0;
    int64_t cnt = 0;
    float ob = 0.f;
    cdlm_status st = cdlm_code_thresh(j->p, j->mult, &th, &cnt);
    if (st == CDLM_OK) st = cdlm_obj_best(j->p, &ob);
    std::vector<float> codes;
    if (st == CDLM_OK && j->keep_codes) {
        codes.resize((size_t)(j->L4 / 4) * j->prm.M * j->N);
        st = cdlm_get(j->p, CDLM_F_XBEST_SLICE, codes.data(), codes.size() * sizeof(float));
    }
    std::string err;
    if (st != CDLM_OK) {
        const char *e = cdlm_last_error(pctx ? pctx : srv->ctx);
        err = e ? e : "";
    }
    cdlm_problem_destroy(j->p);   // the arena is the server's: it goes back to the pool
    j->p = nullptr;
    return_arena(srv, j, st == CDLM_OK && !j->streamed);
    std::lock_guard<std::mutex> lk(srv->mu);
    j->t_done = Clock::now();
    if (st == CDLM_OK) {
        j->thresh = th;
        j->count = cnt;
        j->obj_best = ob;
        j->codes.swap(codes);
        j->state = CDLM_JOB_DONE;
        srv->latencies.push_back(std::chrono::duration<double>(j->t_done - j->t_submit).count());
    } else {
        j->state = CDLM_JOB_FAILED;
        j->error = err;
    }
    j->s.clear();
    j->s.shrink_to_fit();
    srv->cv_done.notify_all();
}

void drop_batch(cdlm_server *srv) {
    if (srv->batch) cdlm_batch_destroy(srv->batch);
    srv->batch = nullptr;
    srv->membership_changed = true;
}

// One tick over the resident jobs; returns false on a device error (all resident jobs fail).
bool tick(cdlm_server *srv) {
    std::vector<Job *> act;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        act = srv->active;
    }
    if (act.empty()) return true;
    std::vector<double> tr(act.size());
    if (act.size() == 1 && act[0]->streamed) {
        // one outer iteration, chunk by chunk through the device window
        drop_batch(srv);
        if (cdlm_optimize(act[0]->p, 1, srv->opts.csc_iter, srv->opts.cdl_iter, tr.data()) != CDLM_OK) return false;
    } else {
    if (srv->membership_changed || !srv->batch) {
        if (srv->batch) cdlm_batch_destroy(srv->batch);
        srv->batch = nullptr;
        std::vector<cdlm_problem *> ps;
        for (Job *j : act) ps.push_back(j->p);
        if (cdlm_batch_create(srv->ctx, ps.data(), (int32_t)ps.size(), &srv->batch) != CDLM_OK) return false;
        cdlm_batch_set_graph(srv->batch, srv->opts.use_graph ? 1 : 0);
        srv->membership_changed = false;
    }
    if (cdlm_batch_optimize(srv->batch, 1, srv->opts.csc_iter, srv->opts.cdl_iter, tr.data()) != CDLM_OK) return false;
    }
    std::vector<Job *> done;
    for (size_t i = 0; i < act.size(); ++i) {
        act[i]->trace.push_back(tr[i]);
        if (++act[i]->outer_done >= act[i]->iter_num) done.push_back(act[i]);
    }
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        ++srv->ticks;
    }
    if (!done.empty()) {
        drop_batch(srv);   // the batch references the finished problems
        for (Job *j : done) finish_job(srv, j);
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->active.erase(std::remove_if(srv->active.begin(), srv->active.end(),
                                         [](Job *j) { return j->state != CDLM_JOB_RUNNING; }),
                          srv->active.end());
    }
    return true;
}

void server_loop(cdlm_server *srv) {
    (void)hipSetDevice(srv->ctx->device);   // arenas are allocated from this thread
    for (;;) {
        std::vector<Job *> adm;
        std::vector<void *> to_free;
        {
            std::unique_lock<std::mutex> lk(srv->mu);
            srv->cv_work.wait(lk, [&] { return srv->stop || !srv->active.empty() || !srv->queue.empty(); });
            if (srv->stop) break;
            adm = pick_admissions(srv, &to_free);
        }
        for (void *a : to_free) (void)hipFree(a);
        for (Job *j : adm) {
            if (j->streamed) {
                cdlm_problem *p = nullptr;
                cdlm_status st = cdlm_problem_create_streamed(srv->ctx, &j->prm, j->s.data(), j->L4, j->N,
                                                              j->D0.empty() ? nullptr : j->D0.data(), nullptr, j->seed,
                                                              j->arena_bytes, &p);
                std::lock_guard<std::mutex> lk(srv->mu);
                if (st != CDLM_OK) {
                    srv->in_use -= j->arena_bytes;
                    j->arena_bytes = 0;
                    j->state = CDLM_JOB_FAILED;
                    j->error = ctx_error(srv);
                    j->t_done = Clock::now();
                    srv->cv_done.notify_all();
                    continue;
                }
                int32_t Lg = 0, gamma = 0;
                cdlm_problem_dims(p, &Lg, &gamma, nullptr, nullptr, nullptr);
                j->p = p;
                j->Lg = Lg;
                j->gamma = gamma;
                j->t_admit = Clock::now();
                srv->active.push_back(j);
                srv->membership_changed = true;
                continue;
            }
            if (!j->arena) {
                if (hipMalloc(&j->arena, j->arena_bytes) != hipSuccess) {
                    (void)hipGetLastError();
                    j->arena = nullptr;
                    return_arena(srv, j, false);
                    fail_job(srv, j, "no device memory for the job's arena");
                    continue;
                }
            }
            cdlm_problem *p = nullptr;
            cdlm_status st = cdlm_problem_create_in(srv->ctx, &j->prm, j->s.data(), j->L4, j->N,
                                                    j->D0.empty() ? nullptr : j->D0.data(), j->seed, j->arena,
                                                    j->arena_bytes, &p);
            if (st != CDLM_OK) {
                return_arena(srv, j, true);
                fail_job(srv, j, ctx_error(srv));
                continue;
            }
            int32_t Lg = 0, gamma = 0;
            cdlm_problem_dims(p, &Lg, &gamma, nullptr, nullptr, nullptr);
            std::lock_guard<std::mutex> lk(srv->mu);
            j->p = p;
            j->Lg = Lg;
            j->gamma = gamma;
            j->t_admit = Clock::now();
            srv->active.push_back(j);   // a zero-iteration job finishes below without a tick
            srv->membership_changed = true;
        }
        // zero-iteration jobs finish at admission
        std::vector<Job *> zero;
        {
            std::lock_guard<std::mutex> lk(srv->mu);
            for (Job *j : srv->active)
                if (j->iter_num == 0 && j->state == CDLM_JOB_RUNNING) zero.push_back(j);
        }
        if (!zero.empty()) {
            drop_batch(srv);
            for (Job *j : zero) finish_job(srv, j);
            std::lock_guard<std::mutex> lk(srv->mu);
            srv->active.erase(std::remove_if(srv->active.begin(), srv->active.end(),
                                             [](Job *j) { return j->state != CDLM_JOB_RUNNING; }),
                              srv->active.end());
        }
        if (!tick(srv)) {
            const std::string err = ctx_error(srv);
            set_error(srv, err);
            drop_batch(srv);
            std::vector<Job *> act;
            {
                std::lock_guard<std::mutex> lk(srv->mu);
                act.swap(srv->active);
            }
            for (Job *j : act) {
                if (j->p) cdlm_problem_destroy(j->p);
                j->p = nullptr;
                return_arena(srv, j, false);
                fail_job(srv, j, err);
            }
        }
    }
    // stop: drop everything still resident or queued
    drop_batch(srv);
    std::vector<Job *> act;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        act.swap(srv->active);
    }
    for (Job *j : act) {
        if (j->p) cdlm_problem_destroy(j->p);
        j->p = nullptr;
        return_arena(srv, j, false);
        fail_job(srv, j, "server stopped");
    }
    std::vector<std::pair<void *, size_t>> pool;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        pool.swap(srv->pool);
        srv->pool_bytes = 0;
    }
    for (auto &a : pool) (void)hipFree(a.first);
    std::lock_guard<std::mutex> lk(srv->mu);
    for (uint64_t id : srv->queue) {
        Job *j = srv->jobs[id].get();
        j->state = CDLM_JOB_FAILED;
        j->error = "server stopped";
    }
    srv->queue.clear();
    srv->cv_done.notify_all();
}

// ---- concurrent mode
// One job on a worker's context: its problem in its arena (allocated here when not recycled),
// its outer iterations one call each (the server may stop between two), then its results.
void run_job(cdlm_server *srv, cdlm_ctx *ctx, Job *j) {
    auto fail = [&](const std::string &why, bool reuse) {
        if (j->p) cdlm_problem_destroy(j->p);
        j->p = nullptr;
        return_arena(srv, j, reuse);
        fail_job(srv, j, why);
    };
    auto stopping = [&] {
        std::lock_guard<std::mutex> lk(srv->mu);
        return srv->stop;
    };
    if (!ctx) return fail("no device context for the job", false);
    if (stopping()) return fail("server stopped", !j->streamed);
    cdlm_problem *p = nullptr;
    cdlm_status st;
    if (j->streamed) {
        st = cdlm_problem_create_streamed(ctx, &j->prm, j->s.data(), j->L4, j->N, j->D0.empty() ? nullptr : j->D0.data(),
                                          nullptr, j->seed, j->arena_bytes, &p);
    } else {
        if (!j->arena && hipMalloc(&j->arena, j->arena_bytes) != hipSuccess) {
            (void)hipGetLastError();
            j->arena = nullptr;
            return fail("no device memory for the job's arena", false);
        }
        st = cdlm_problem_create_in(ctx, &j->prm, j->s.data(), j->L4, j->N, j->D0.empty() ? nullptr : j->D0.data(),
                                    j->seed, j->arena, j->arena_bytes, &p);
    }
    if (st != CDLM_OK) {
        const char *e = cdlm_last_error(ctx);
        return fail(e ? e : "problem creation failed", !j->streamed);
    }
    int32_t Lg = 0, gamma = 0;
    cdlm_problem_dims(p, &Lg, &gamma, nullptr, nullptr, nullptr);
    j->p = p;
    j->Lg = Lg;
    j->gamma = gamma;
    for (int it = 0; it < j->iter_num; ++it) {
        if (stopping()) return fail("server stopped", false);
        double tr = 0.0;
        if (cdlm_optimize(p, 1, srv->opts.csc_iter, srv->opts.cdl_iter, &tr) != CDLM_OK) {
            const char *e = cdlm_last_error(ctx);
            return fail(e ? e : "device error", false);
        }
        j->trace.push_back(tr);
        std::lock_guard<std::mutex> lk(srv->mu);
        ++srv->ticks;
    }
    finish_job(srv, j, ctx);
}

void worker_loop(cdlm_server *srv) {
    cdlm_ctx *ctx = nullptr;
    if (cdlm_create(&ctx, srv->ctx->device) != CDLM_OK) ctx = nullptr;   // its jobs then fail
    for (;;) {
        Job *j = nullptr;
        {
            std::unique_lock<std::mutex> lk(srv->mu);
            srv->cv_ready.wait(lk, [&] { return srv->stop || !srv->ready.empty(); });
            if (srv->ready.empty()) break;   // stopping, nothing handed over
            j = srv->ready.front();
            srv->ready.pop_front();
        }
        run_job(srv, ctx, j);
        {
            std::lock_guard<std::mutex> lk(srv->mu);
            srv->active.erase(std::remove(srv->active.begin(), srv->active.end(), j), srv->active.end());
            ++srv->idle;
            ++srv->gen;
        }
        srv->cv_work.notify_all();
    }
    if (ctx) cdlm_destroy(ctx);
}

// Admits queued jobs whenever a worker is idle and a submission or a job leaving the device may
// have made room; the workers do the rest.
void dispatch_loop(cdlm_server *srv) {
    (void)hipSetDevice(srv->ctx->device);   // pooled arenas are freed from this thread
    uint64_t seen = ~0ull;
    for (;;) {
        std::vector<Job *> adm;
        std::vector<void *> to_free;
        {
            std::unique_lock<std::mutex> lk(srv->mu);
            srv->cv_work.wait(lk, [&] { return srv->stop || (!srv->queue.empty() && srv->idle > 0 && srv->gen != seen); });
            if (srv->stop) break;
            seen = srv->gen;
            adm = pick_admissions(srv, &to_free);
            for (Job *j : adm) {
                j->t_admit = Clock::now();
                srv->active.push_back(j);
                --srv->idle;
            }
        }
        for (void *a : to_free) (void)hipFree(a);   // before the workers allocate in their place
        if (!adm.empty()) {
            std::lock_guard<std::mutex> lk(srv->mu);
            for (Job *j : adm) srv->ready.push_back(j);
            ++srv->gen;   // more may fit once these are placed
        }
        srv->cv_ready.notify_all();
    }
    srv->cv_ready.notify_all();
    for (std::thread &t : srv->workers)
        if (t.joinable()) t.join();
    std::vector<std::pair<void *, size_t>> pool;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        pool.swap(srv->pool);
        srv->pool_bytes = 0;
    }
    for (auto &a : pool) (void)hipFree(a.first);
    std::lock_guard<std::mutex> lk(srv->mu);
    for (uint64_t id : srv->queue) {
        Job *j = srv->jobs[id].get();
        j->state = CDLM_JOB_FAILED;
        j->error = "server stopped";
    }
    srv->queue.clear();
    srv->cv_done.notify_all();
}

double quantile(std::vector<double> v, double q) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    // nearest rank
    size_t k = (size_t)std::ceil(q * (double)v.size());
    if (k < 1) k = 1;
    if (k > v.size()) k = v.size();
    return v[k - 1];
}

}  // namespace

extern "C" cdlm_status cdlm_server_create(int32_t device, const cdlm_server_opts *opts, cdlm_server **out) {
    if (!out) return CDLM_ERR_ARG;
    *out = nullptr;
    std::unique_ptr<cdlm_server> srv(new cdlm_server());
    cdlm_status st = cdlm_create(&srv->ctx, device);
    if (st != CDLM_OK) return st;
    if (opts) srv->opts = *opts;
    if (srv->opts.csc_iter <= 0) srv->opts.csc_iter = 15;   // WRAPPER_setup.jl:8
    if (srv->opts.cdl_iter <= 0) srv->opts.cdl_iter = 1;    // WRAPPER_setup.jl:9
    size_t free_b = 0, total_b = 0;
    if ((st = cdlm_mem_info(srv->ctx, &free_b, &total_b)) != CDLM_OK) {
        cdlm_destroy(srv->ctx);
        return st;
    }
    srv->budget = srv->opts.mem_budget ? (size_t)srv->opts.mem_budget : (size_t)(0.8 * (double)free_b);
    if (srv->budget > free_b) srv->budget = free_b;
    if (srv->opts.fill_seqs > 0) {
        srv->fill = srv->opts.fill_seqs;
    } else if (srv->opts.fill_seqs == 0) {
        // three resident work-groups per CU for the CSC kernel, four times over
        hipDeviceProp_t prop;
        int cus = 256;
        if (hipGetDeviceProperties(&prop, device) == hipSuccess) cus = prop.multiProcessorCount;
        srv->fill = 12L * cus;
    } else {
        srv->fill = 0;   // no limit
    }
    cdlm_server *raw = srv.get();
    if (raw->opts.lockstep) {
        raw->th = std::thread(server_loop, raw);
    } else {
        // four workers (the device's hardware queues per process), or max_active when smaller
        const int nw = raw->opts.max_active > 0 ? std::min(raw->opts.max_active, 4) : 4;
        raw->idle = nw;
        for (int i = 0; i < nw; ++i) raw->workers.emplace_back(worker_loop, raw);
        raw->th = std::thread(dispatch_loop, raw);
    }
    *out = srv.release();
    return CDLM_OK;
}

extern "C" cdlm_status cdlm_server_destroy(cdlm_server *srv) {
    if (!srv) return CDLM_OK;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->stop = true;
    }
    srv->cv_work.notify_all();
    srv->cv_ready.notify_all();
    if (srv->th.joinable()) srv->th.join();   // the dispatcher joins the workers
    cdlm_destroy(srv->ctx);
    delete srv;
    return CDLM_OK;
}

extern "C" const char *cdlm_server_last_error(cdlm_server *srv) {
    if (!srv) return "null server";
    // The server thread may replace last_error at any time, so the caller gets a copy owned
    // by its own thread, valid until that thread calls this function again.
    static thread_local std::string copy;
    std::lock_guard<std::mutex> lk(srv->mu);
    copy = srv->last_error;
    return copy.c_str();
}

extern "C" cdlm_status cdlm_server_submit(cdlm_server *srv, const cdlm_job_spec *spec, uint64_t *id) {
    if (!srv || !spec || !id) return CDLM_ERR_ARG;
    auto reject = [&](cdlm_status st, const std::string &m) {
        set_error(srv, m);
        return st;
    };
    const cdlm_params &prm = spec->params;
    if (!spec->s || !prm.filter_len || prm.M <= 0 || spec->N <= 0 || spec->L4 <= 0 || spec->L4 % 4 || spec->iter_num < 0)
        return reject(CDLM_ERR_ARG, "bad job spec");
    std::unique_ptr<Job> j(new Job());
    j->fl.assign(prm.filter_len, prm.filter_len + prm.M);
    int gamma = 0;
    for (int f : j->fl) {
        if (f <= 0 || f % 4) return reject(CDLM_ERR_ARG, "filter lengths must be positive multiples of 4");
        gamma = std::max(gamma, f);
    }
    if (gamma > 32 || spec->L4 + 2 * gamma > 2048)
        return reject(CDLM_ERR_UNSUPPORTED, "job shape outside the kernels' range (filter length <= 32, Lg <= 2048)");
    j->footprint = cdlm_footprint_bytes(spec->L4, spec->N, prm.M, gamma, 0);
    j->work = (double)spec->iter_num * spec->N * (double)(spec->L4 + 2 * gamma) * prm.M;
    // a job beyond the budget is streamed through a device window of the budget (alone on the
    // device); its state then lives in pinned host memory
    j->streamed = j->footprint > srv->budget;
    if (j->streamed && cdlm_stream_min_bytes(spec->L4, spec->N, prm.M, gamma) > srv->budget) {
        char buf[256];
        snprintf(buf, sizeof(buf), "job needs %zu device bytes and at least %zu streamed, more than the server budget of %zu",
                 j->footprint, cdlm_stream_min_bytes(spec->L4, spec->N, prm.M, gamma), srv->budget);
        return reject(CDLM_ERR_OOM, buf);
    }
    j->prm = prm;
    j->prm.filter_len = j->fl.data();
    j->L4 = spec->L4;
    j->N = spec->N;
    j->s.assign(spec->s, spec->s + (size_t)spec->L4 * spec->N);
    const int Lg = spec->L4 + 2 * gamma;
    if (spec->D0) j->D0.assign(spec->D0, spec->D0 + (size_t)Lg * prm.M);
    j->seed = spec->seed;
    j->iter_num = spec->iter_num;
    j->mult = spec->thresh_mult;
    j->keep_codes = spec->keep_codes;
    j->t_submit = Clock::now();
    std::lock_guard<std::mutex> lk(srv->mu);
    if (srv->stop) return CDLM_ERR_STATE;
    j->id = srv->next_id++;
    *id = j->id;
    srv->queue.push_back(j->id);
    srv->jobs[j->id] = std::move(j);
    ++srv->gen;
    srv->cv_work.notify_all();
    return CDLM_OK;
}

extern "C" cdlm_status cdlm_server_poll(cdlm_server *srv, uint64_t id, int32_t *state) {
    if (!srv || !state) return CDLM_ERR_ARG;
    std::lock_guard<std::mutex> lk(srv->mu);
    auto it = srv->jobs.find(id);
    if (it == srv->jobs.end()) return CDLM_ERR_ARG;
    *state = it->second->state;
    return CDLM_OK;
}

extern "C" cdlm_status cdlm_server_wait(cdlm_server *srv, uint64_t id, int32_t timeout_ms, int32_t *state) {
    if (!srv) return CDLM_ERR_ARG;
    std::unique_lock<std::mutex> lk(srv->mu);
    auto it = srv->jobs.find(id);
    if (it == srv->jobs.end()) return CDLM_ERR_ARG;
    Job *j = it->second.get();
    auto done = [&] { return j->state == CDLM_JOB_DONE || j->state == CDLM_JOB_FAILED; };
    if (timeout_ms < 0) srv->cv_done.wait(lk, done);
    else srv->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), done);
    if (state) *state = j->state;
    return CDLM_OK;
}

extern "C" cdlm_status cdlm_server_result(cdlm_server *srv, uint64_t id, cdlm_job_result *res, double *obj_trace,
                                          float *codes, size_t codes_bytes) {
    if (!srv || !res) return CDLM_ERR_ARG;
    std::lock_guard<std::mutex> lk(srv->mu);
    auto it = srv->jobs.find(id);
    if (it == srv->jobs.end()) return CDLM_ERR_ARG;
    Job *j = it->second.get();
    if (j->state == CDLM_JOB_FAILED) {
        srv->last_error = j->error;
        return CDLM_ERR_STATE;
    }
    if (j->state != CDLM_JOB_DONE) return CDLM_ERR_STATE;
    res->latency_s = std::chrono::duration<double>(j->t_done - j->t_submit).count();
    res->queue_s = std::chrono::duration<double>(j->t_admit - j->t_submit).count();
    res->service_s = std::chrono::duration<double>(j->t_done - j->t_admit).count();
    res->thresh = j->thresh;
    res->count = j->count;
    res->obj_best = j->obj_best;
    res->iter_num = j->iter_num;
    res->L = j->L4 / 4;
    res->M = j->prm.M;
    res->N = j->N;
    if (obj_trace && !j->trace.empty()) memcpy(obj_trace, j->trace.data(), j->trace.size() * sizeof(double));
    if (codes) {
        if (j->codes.empty() || codes_bytes != j->codes.size() * sizeof(float)) {
            srv->last_error = "codes were not kept for this job or the buffer size differs";
            return CDLM_ERR_ARG;
        }
        memcpy(codes, j->codes.data(), codes_bytes);
    }
    return CDLM_OK;
}

extern "C" cdlm_status cdlm_server_release(cdlm_server *srv, uint64_t id) {
    if (!srv) return CDLM_ERR_ARG;
    std::lock_guard<std::mutex> lk(srv->mu);
    auto it = srv->jobs.find(id);
    if (it == srv->jobs.end()) return CDLM_ERR_ARG;
    if (it->second->state != CDLM_JOB_DONE && it->second->state != CDLM_JOB_FAILED) return CDLM_ERR_STATE;
    srv->jobs.erase(it);
    return CDLM_OK;
}

extern "C" cdlm_status cdlm_server_stats(cdlm_server *srv, cdlm_server_stats_t *out) {
    if (!srv || !out) return CDLM_ERR_ARG;
    std::lock_guard<std::mutex> lk(srv->mu);
    out->completed = (int64_t)srv->latencies.size();
    out->active = (int64_t)srv->active.size();
    out->queued = (int64_t)srv->queue.size();
    out->ticks = srv->ticks;
    out->p50_s = quantile(srv->latencies, 0.50);
    out->p99_s = quantile(srv->latencies, 0.99);
    out->bytes_in_use = srv->in_use + srv->pool_bytes;   // resident jobs' arenas and the pool
    out->budget_bytes = srv->budget;
    return CDLM_OK;
}
