// serve.hip -- the continuous-batching server of the C ABI (voss_serve_*).
//
// A scheduler thread owns the context while the server runs.  Requests
// (an interval, or a batch of Miller-Rabin candidates) wait in arrival order
// and are admitted while a request slot and the HBM budget are free.
//
// Interval requests.  Each is planned into LDS windows once, at admission,
// and gets a slot of per-request accumulators in HBM (JobAcc) that stays on
// the device until the request completes.  A step is one executor plan made
// of up to windows_per_step windows taken round-robin from every active
// request (bucket fill, sieve, stitch per round), followed on the same
// stream by a kernel that adds the step's sums into the requests' slots.
// Steps are enqueued without waiting for the device: there are two sets of
// step buffers, so while the device runs step k the host plans and enqueues
// step k+1, and the results of step k (pairs across step edges, large-gap
// candidates, overflow flags, finished requests) are collected when its
// event has completed -- the device always has the next step queued behind
// the running one, and no step waits for the host.
//
// Miller-Rabin requests.  Admitted batches are packed into one of two
// pinned staging sets, copied and tested on a second stream (default
// priority) beside the sieve steps, and completed from their event.
//
// Results are exactly those of the request run alone: integer sums do not
// depend on order, and the pairs across step edges are joined on the host in
// window order.  A step whose bucket rooms or large-gap list overflowed
// (which the room sizing makes essentially impossible) restarts the
// requests it touched from their first window, with a larger large-gap list.
#include <condition_variable>
#include <deque>
#include <memory>
#include <thread>
#include <unordered_map>

#include "engine.h"

namespace {

struct ServeReq {
    uint64_t id = 0;
    bool is_mr = false;
    uint64_t lo = 0, hi = 0;
    std::vector<uint64_t> cand;
    std::vector<uint8_t> verdict;
    double t_submit = 0, t_done = 0;
    uint64_t hbm = 0;
    // interval requests
    DescVec win;         // planned windows, job = slot
    size_t next = 0;     // windows handed to steps so far
    int slot = -1;
    uint32_t epoch = 0;  // bumped by a restart; steps planned before it are ignored
    bool needs_init = true;  // the slot's accumulators must be cleared first
    JobHost J;
    uint64_t carry_last = 0;
    bool carry_has = false;
    voss_interval_result res{};
    std::vector<voss_large_gap> large;
    int status = VOSS_OK;
    bool done = false;
};

using ReqPtr = std::shared_ptr<ServeReq>;

// One in-flight sieve step: its buffers, plan, and which request each slot
// held (and at which epoch) when it was planned.
struct Step {
    ExecBufs B;
    std::unique_ptr<Executor> X;
    hipEvent_t start = nullptr, done = nullptr;
    std::vector<ReqPtr> by_slot;
    std::vector<uint32_t> epoch_of_slot;
    std::vector<ReqPtr> finishing;  // requests whose last windows are in this step
    bool busy = false;
    uint64_t seq = 0;
};

// One in-flight Miller-Rabin launch over the packed candidates of its batches.
struct MrLaunch {
    HostBuf<uint64_t> h_n;
    HostBuf<uint8_t> h_v;
    DevBuf<uint64_t> d_n;
    DevBuf<uint8_t> d_v;
    DevBuf<uint32_t> d_surv;
    DevBuf<uint32_t> d_nsurv;
    hipEvent_t done = nullptr;
    std::vector<ReqPtr> reqs;
    uint64_t n = 0;
    bool busy = false;
};

// The buffers, events and MR stream of the scheduler.  They belong to the
// context and outlive a server, so a later server on the same context
// starts with buffers that earlier requests have grown (growing a device
// or pinned buffer synchronises the device).
struct ServeCache {
    Step steps[2];
    MrLaunch mr[2];
    DevBuf<JobAcc> d_acc;   // per-slot accumulators, resident while requests run
    HostBuf<JobAcc> h_acc;  // read-back of finished slots
    hipStream_t mr_stream = nullptr;
};

void free_serve_cache(void* p) {
    ServeCache* k = static_cast<ServeCache*>(p);
    for (Step& S : k->steps) {
        S.X.reset();
        S.B.release();
        if (S.start) (void)hipEventDestroy(S.start);
        if (S.done) (void)hipEventDestroy(S.done);
    }
    for (MrLaunch& M : k->mr) {
        M.h_n.release();
        M.h_v.release();
        M.d_n.release();
        M.d_v.release();
        M.d_surv.release();
        M.d_nsurv.release();
        if (M.done) (void)hipEventDestroy(M.done);
    }
    if (k->mr_stream) (void)hipStreamDestroy(k->mr_stream);
    k->d_acc.release();
    k->h_acc.release();
    delete k;
}

}  // namespace

struct voss_server {
    voss_ctx* c = nullptr;
    voss_serve_config cfg{};
    std::thread th;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<ReqPtr> queue;
    std::unordered_map<uint64_t, ReqPtr> all;
    uint64_t next_id = 1;
    bool stopping = false;
    int err = VOSS_OK;
    std::string errmsg;
    // owned by the scheduler thread
    std::vector<ReqPtr> active;
    std::vector<int> free_slots;
    uint64_t hbm_used = 0;
    ServeCache* k = nullptr;  // the context's scheduler buffers
    uint64_t step_seq = 0;
    uint64_t lcap = 0;      // large-gap list capacity of a step (grows on overflow)
    bool rooms_exact = false;  // set by a bucket overflow: exact rooms from then on
    // statistics (voss_serve_stats)
    std::atomic<uint64_t> arrivals{0};  // submissions so far (wakes the collector)
    std::atomic<uint64_t> n_steps{0};
    std::atomic<uint64_t> n_windows{0};
    double busy_ms = 0;      // device time of the sieve steps
    double t_start = 0;
    std::mutex stat_mu;
};

namespace {

constexpr uint64_t kSlotBytes = sizeof(JobAcc);

uint64_t req_hbm(const voss_server* s, const ServeReq& r) {
    if (r.is_mr) return r.cand.size() * (8 + 1 + 4);
    return kSlotBytes + (uint64_t)s->cfg.large_capacity * sizeof(LargeRaw);
}

void serve_fail_all(voss_server* s, int code, const std::string& msg) {
    std::lock_guard<std::mutex> lk(s->mu);
    s->err = code;
    s->errmsg = msg;
    for (auto& kv : s->all)
        if (!kv.second->done) {
            kv.second->status = code;
            kv.second->done = true;
        }
    s->cv_done.notify_all();
}

void complete(voss_server* s, const std::vector<ReqPtr>& finished) {
    if (finished.empty()) return;
    const double t = now_s();
    std::lock_guard<std::mutex> lk(s->mu);
    for (auto& r : finished) {
        r->t_done = t;
        r->done = true;
        s->hbm_used -= r->hbm;
        if (!r->is_mr) s->free_slots.push_back(r->slot);
    }
    s->cv_done.notify_all();
}

// Starts a request over from its first window (after an overflow).
void restart(ServeReq* r) {
    r->epoch++;
    r->next = 0;
    r->needs_init = true;
    r->J = JobHost();
    r->carry_last = 0;
    r->carry_has = false;
}

// Plans and enqueues one step into S: up to windows_per_step windows
// round-robin from the active requests; the slot accumulators of newly
// admitted (or restarted) requests are cleared first.  *enqueued = false
// when no request has windows left to hand out.
int enqueue_step(voss_server* s, Step& S, bool* enqueued) {
    voss_ctx* c = s->c;
    *enqueued = false;
    size_t want = 0;
    for (auto& r : s->active)
        if (r->next < r->win.size()) want++;
    if (!want) return VOSS_OK;
    const size_t quota = std::max<size_t>(1, s->cfg.windows_per_step / want);
    S.X.reset(new Executor(c, &S.B));
    Executor& X = *S.X;
    X.P.njobs = s->cfg.max_active;
    X.copy_acc = false;
    X.lcap = s->lcap;
    S.by_slot.assign(s->cfg.max_active, nullptr);
    S.epoch_of_slot.assign(s->cfg.max_active, 0);
    S.finishing.clear();
    for (auto& r : s->active) {
        if (r->needs_init) {
            HIPCHK(launch_init_acc(s->k->d_acc.p + r->slot, 1, c->stream));
            r->needs_init = false;
        }
        const size_t left = r->win.size() - r->next;
        if (!left) continue;
        const size_t k = std::min(left, quota);
        X.P.desc.insert(X.P.desc.end(), r->win.begin() + (long)r->next,
                        r->win.begin() + (long)(r->next + k));
        r->next += k;
        S.by_slot[r->slot] = r;
        S.epoch_of_slot[r->slot] = r->epoch;
        if (r->next == r->win.size()) S.finishing.push_back(r);
    }
    X.plan_rounds();
    int rc = X.stage();
    if (rc) return rc;
    HIPCHK(hipEventRecord(S.start, c->stream));
    // after a bucket overflow the server sizes every step's buckets exactly
    // (a count pass and a host prefix sum before the fill: the host waits
    // for that step's count, the only wait inside a step)
    if ((rc = s->rooms_exact ? X.enqueue_exact(false) : X.enqueue_direct(false))) return rc;
    // the step's sums into the slots (stream order puts this after the
    // step and before any later step or read of a slot)
    HIPCHK(launch_add_acc(s->k->d_acc.p, S.B.d_acc.p, s->cfg.max_active, c->stream));
    HIPCHK(hipEventRecord(S.done, c->stream));
    S.busy = true;
    S.seq = s->step_seq++;
    s->n_steps++;
    s->n_windows += X.P.desc.size();
    *enqueued = true;
    return VOSS_OK;
}

// Collects a finished step: overflow check, large-gap candidates, pairs
// across its round edges and the previous steps, finished requests.
int collect_step(voss_server* s, Step& S) {
    voss_ctx* c = s->c;
    HIPCHK(hipEventSynchronize(S.done));
    S.busy = false;
    float ms = 0;
    if (hipEventElapsedTime(&ms, S.start, S.done) == hipSuccess) {
        std::lock_guard<std::mutex> g(s->stat_mu);
        s->busy_ms += ms;
    }
    Executor& X = *S.X;
    const unsigned long long nlarge = S.B.h_flags.p[0];
    auto current = [&](uint32_t slot) -> ServeReq* {
        ServeReq* r = S.by_slot[slot].get();
        return r && r->epoch == S.epoch_of_slot[slot] ? r : nullptr;
    };
    if (S.B.h_flags.p[1] || nlarge > X.lcap) {
        // a bucket outgrew its room or the large list its capacity: the
        // requests of this step start over (with exact rooms, or a larger
        // list), so a restart cannot overflow the same way again
        if (S.B.h_flags.p[1]) s->rooms_exact = true;
        if (nlarge > X.lcap) s->lcap = std::max<uint64_t>(s->lcap, nlarge + (nlarge >> 2));
        for (uint32_t k = 0; k < s->cfg.max_active; k++)
            if (ServeReq* r = current(k)) restart(r);
        return VOSS_OK;
    }
    if (nlarge) {
        HIPCHK(S.B.h_large.reserve(nlarge));
        HIPCHK(hipMemcpy(S.B.h_large.p, S.B.d_large.p, nlarge * sizeof(LargeRaw),
                         hipMemcpyDeviceToHost));
        for (uint64_t k = 0; k < nlarge; k++) {
            const LargeRaw& x = S.B.h_large.p[k];
            if (ServeReq* r = current(x.job))
                r->J.large.push_back({x.p_before + x.gap, (int32_t)x.gap, merit_of(x.p_before, x.gap)});
        }
    }
    // edges between each round and what the request saw before it
    const SegRec* rec = S.B.h_rec.p;
    std::vector<uint8_t> seen(s->cfg.max_active, 0);
    for (const Round& R : X.rounds) {
        std::fill(seen.begin(), seen.end(), 0);
        for (size_t i = R.off; i < R.off + R.nd; i++) {
            if (rec[i].count == 0) continue;
            const uint32_t slot = X.P.desc[i].job;
            ServeReq* r = current(slot);
            if (!r) continue;
            if (!seen[slot]) {
                seen[slot] = 1;
                if (r->carry_has) host_pair(r->J, r->carry_last, rec[i].first, true);
            }
            r->carry_has = true;
            r->carry_last = rec[i].last;
        }
    }
    // requests whose last windows were in this step: their slot is final
    std::vector<ReqPtr> finished;
    for (auto& r : S.finishing) {
        if (r->epoch != S.epoch_of_slot[r->slot] || r->next != r->win.size()) continue;
        HIPCHK(s->k->h_acc.reserve(1));
        HIPCHK(hipMemcpy(s->k->h_acc.p, s->k->d_acc.p + r->slot, sizeof(JobAcc), hipMemcpyDeviceToHost));
        add_acc(r->J, *s->k->h_acc.p);
        r->large.resize(s->cfg.large_capacity);
        r->res.large = r->large.data();
        r->res.large_capacity = s->cfg.large_capacity;
        finish_interval(r->J, r->lo, r->hi, &r->res);
        r->large.resize(std::min<uint64_t>(r->res.large_count, s->cfg.large_capacity));
        if (r->J.overflow) r->status = VOSS_ERANGE;
        finished.push_back(r);
    }
    for (auto& r : finished)
        s->active.erase(std::find(s->active.begin(), s->active.end(), r));
    complete(s, finished);
    (void)c;
    return VOSS_OK;
}

// Packs admitted MR batches into M and launches them on the MR stream.
int launch_mr_batches(voss_server* s, MrLaunch& M, std::vector<ReqPtr>& batches) {
    voss_ctx* c = s->c;
    uint64_t n = 0;
    for (auto& r : batches) n += r->cand.size();
    M.reqs.swap(batches);
    M.n = n;
    M.busy = true;
    if (n == 0) return VOSS_OK;
    HIPCHK(M.h_n.reserve(n));
    HIPCHK(M.h_v.reserve(n));
    HIPCHK(M.d_n.reserve(n));
    HIPCHK(M.d_v.reserve(n));
    HIPCHK(M.d_surv.reserve(2 * n));
    HIPCHK(M.d_nsurv.reserve(3));
    uint64_t o = 0;
    for (auto& r : M.reqs) {
        std::memcpy(M.h_n.p + o, r->cand.data(), r->cand.size() * 8);
        o += r->cand.size();
    }
    HIPCHK(hipMemcpyAsync(M.d_n.p, M.h_n.p, n * 8, hipMemcpyHostToDevice, s->k->mr_stream));
    HIPCHK(launch_mr(M.d_n.p, n, M.d_v.p, M.d_surv.p, M.d_nsurv.p, c->num_cus, s->k->mr_stream));
    HIPCHK(hipMemcpyAsync(M.h_v.p, M.d_v.p, n, hipMemcpyDeviceToHost, s->k->mr_stream));
    HIPCHK(hipEventRecord(M.done, s->k->mr_stream));
    return VOSS_OK;
}

// Completes an MR launch whose event has finished (wait: block for it).
int collect_mr(voss_server* s, MrLaunch& M, bool wait, bool* collected) {
    *collected = false;
    if (!M.busy) return VOSS_OK;
    if (M.n) {
        if (wait) {
            HIPCHK(hipEventSynchronize(M.done));
        } else {
            const hipError_t q = hipEventQuery(M.done);
            if (q == hipErrorNotReady) return VOSS_OK;
            HIPCHK(q);
        }
    }
    uint64_t o = 0;
    for (auto& r : M.reqs) {
        r->verdict.assign(M.h_v.p + o, M.h_v.p + o + r->cand.size());
        o += r->cand.size();
    }
    M.busy = false;
    complete(s, M.reqs);
    M.reqs.clear();
    *collected = true;
    return VOSS_OK;
}

// The scheduler: admission, step and MR launches, collection.  The step
// that was enqueued first is always collected first.
void serve_loop(voss_server* s) {
    voss_ctx* c = s->c;
    t_scheduler = true;
    if (hipSetDevice(c->device) != hipSuccess) {
        serve_fail_all(s, VOSS_EDEVICE, "hipSetDevice");
        return;
    }
    const uint32_t W = c->tune.window_bytes;
    std::deque<ReqPtr> mr_waiting;  // admitted MR batches not launched yet
    while (true) {
        std::vector<ReqPtr> admitted;
        const bool in_flight = s->k->steps[0].busy || s->k->steps[1].busy || s->k->mr[0].busy ||
                               s->k->mr[1].busy;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            // with nothing on the device, every active request has windows
            // left to hand out (its last step was collected otherwise)
            if (!in_flight && mr_waiting.empty())
                s->cv_work.wait(lk, [&] {
                    return s->stopping || !s->queue.empty() || !s->active.empty();
                });
            if (s->queue.empty() && s->active.empty() && s->stopping && !in_flight &&
                mr_waiting.empty())
                break;
            // admission by slots and HBM budget, in arrival order
            while (!s->queue.empty()) {
                ReqPtr r = s->queue.front();
                const uint64_t cost = req_hbm(s, *r);
                const bool fits = s->hbm_used + cost <= s->cfg.hbm_budget_bytes ||
                                  (s->active.empty() && admitted.empty() && !in_flight);
                if (!fits) break;
                if (!r->is_mr) {
                    if (s->free_slots.empty()) break;
                    r->slot = s->free_slots.back();
                    s->free_slots.pop_back();
                }
                r->hbm = cost;
                s->hbm_used += cost;
                admitted.push_back(r);
                s->queue.pop_front();
            }
        }
        // the device work of this pass holds the process-wide call gate, so
        // no other thread's call runs beside it (it waits for one to finish)
        struct PassGate {
            PassGate() { gate().enter(); }
            ~PassGate() { gate().leave(); }
        } pass_gate;
        int rc = VOSS_OK;
        std::vector<ReqPtr> empty_done;
        for (auto& r : admitted) {
            if (r->is_mr) {
                mr_waiting.push_back(r);
                continue;
            }
            if (r->hi > 7) {
                if ((rc = ensure_primes(c, r->hi / 30 * 30 + 30ull * W + 30))) break;
                Plan P1;
                P1.njobs = 1;
                const uint64_t a = std::max<uint64_t>(r->lo, 7), b = r->hi - 1;
                plan_range(c, P1, a, b, (uint32_t)r->slot, 0, W, (a / 30) & ~4095ull);
                r->win.swap(P1.desc);
            }
            if (r->win.empty()) {
                // no wheel numbers to sieve: the result is the head primes alone
                r->large.resize(s->cfg.large_capacity);
                r->res.large = r->large.data();
                r->res.large_capacity = s->cfg.large_capacity;
                finish_interval(r->J, r->lo, r->hi, &r->res);
                r->large.resize(std::min<uint64_t>(r->res.large_count, s->cfg.large_capacity));
                empty_done.push_back(r);
            } else {
                s->active.push_back(r);
            }
        }
        complete(s, empty_done);
        // Miller-Rabin launches: up to mr_per_step candidates each, two in flight
        for (int k = 0; k < 2 && rc == VOSS_OK && !mr_waiting.empty(); k++) {
            MrLaunch& M = s->k->mr[k];
            if (M.busy) continue;
            std::vector<ReqPtr> pack;
            uint64_t n = 0;
            while (!mr_waiting.empty() &&
                   (pack.empty() || n + mr_waiting.front()->cand.size() <= s->cfg.mr_per_step)) {
                n += mr_waiting.front()->cand.size();
                pack.push_back(mr_waiting.front());
                mr_waiting.pop_front();
            }
            rc = launch_mr_batches(s, M, pack);
        }
        // sieve steps: keep one queued behind the running one
        for (int k = 0; k < 2 && rc == VOSS_OK; k++) {
            Step& S = s->k->steps[s->step_seq & 1];
            if (S.busy) break;
            bool enq = false;
            if ((rc = enqueue_step(s, S, &enq)) || !enq) break;
        }
        if (rc == VOSS_OK && !s->active.empty() && !s->k->steps[0].busy && !s->k->steps[1].busy)
            rc = fail(VOSS_EDEVICE, "serving: active requests without pending windows");
        // collect: poll the oldest step and the MR launches (event queries,
        // a short sleep between rounds) until one of them completes or new
        // requests arrive, which go to admission at once (an MR batch can
        // start while the sieve steps run)
        const uint64_t seen = s->arrivals.load();
        while (rc == VOSS_OK) {
            Step* oldest = nullptr;
            for (Step& S : s->k->steps)
                if (S.busy && (!oldest || S.seq < oldest->seq)) oldest = &S;
            bool progressed = false;
            if (oldest) {
                const hipError_t q = hipEventQuery(oldest->done);
                if (q == hipSuccess) {
                    rc = collect_step(s, *oldest);
                    progressed = true;
                } else if (q != hipErrorNotReady) {
                    rc = hip_fail(q, "serving step");
                }
            }
            for (int k = 0; k < 2 && rc == VOSS_OK; k++) {
                bool got = false;
                rc = collect_mr(s, s->k->mr[k], false, &got);
                progressed |= got;
            }
            const bool idle = !oldest && !s->k->mr[0].busy && !s->k->mr[1].busy;
            if (rc || progressed || idle || s->arrivals.load() != seen) break;
            std::this_thread::sleep_for(std::chrono::microseconds(20));
        }
        if (rc != VOSS_OK) {
            serve_fail_all(s, rc, voss_last_error());
            return;
        }
    }
}

int serve_result(voss_server* s, ReqPtr r, voss_interval_result* out, uint8_t* verdict,
                 double* latency_s) {
    if (latency_s) *latency_s = r->t_done - r->t_submit;
    if (r->status != VOSS_OK) return fail(r->status, s->errmsg.empty() ? "request failed" : s->errmsg);
    if (r->is_mr) {
        if (verdict && !r->verdict.empty()) std::memcpy(verdict, r->verdict.data(), r->verdict.size());
    } else if (out) {
        voss_large_gap* lg = out->large;
        const uint64_t cap = out->large_capacity;
        *out = r->res;
        out->large = lg;
        out->large_capacity = cap;
        if (lg && cap) {
            const size_t n = std::min<size_t>(r->large.size(), cap);
            std::memcpy(lg, r->large.data(), n * sizeof(voss_large_gap));
        }
    }
    s->all.erase(r->id);
    return VOSS_OK;
}

}  // namespace

extern "C" {

int voss_serve_start(voss_ctx* c, const voss_serve_config* cfg, voss_server** out) {
    if (!c || !out) return fail(VOSS_EINVAL, "NULL argument");
    *out = nullptr;
    VOSS_ADMIT(c);  // also refuses a context that already has a server
    int rc = set_device(c);
    if (rc) return rc;
    voss_server* s = new voss_server();
    s->c = c;
    voss_serve_config d{64, 4096, 8ull << 30, 64, 1u << 22};
    if (cfg) {
        if (cfg->max_active) d.max_active = cfg->max_active;
        if (cfg->windows_per_step) d.windows_per_step = cfg->windows_per_step;
        if (cfg->hbm_budget_bytes) d.hbm_budget_bytes = cfg->hbm_budget_bytes;
        if (cfg->large_capacity) d.large_capacity = cfg->large_capacity;
        if (cfg->mr_per_step) d.mr_per_step = cfg->mr_per_step;
    }
    s->cfg = d;
    if (!c->serve_cache) {
        ServeCache* k = new ServeCache();
        hipError_t e = hipStreamCreateWithFlags(&k->mr_stream, hipStreamNonBlocking);
        for (Step& S : k->steps) {
            if (e == hipSuccess) e = hipEventCreate(&S.start);
            if (e == hipSuccess) e = hipEventCreate(&S.done);
        }
        for (MrLaunch& M : k->mr)
            if (e == hipSuccess) e = hipEventCreateWithFlags(&M.done, hipEventDisableTiming);
        if (e != hipSuccess) {
            free_serve_cache(k);
            delete s;
            return hip_fail(e, "serve setup");
        }
        c->serve_cache = k;
        c->free_serve_cache = free_serve_cache;
    }
    s->k = static_cast<ServeCache*>(c->serve_cache);
    const hipError_t e = s->k->d_acc.reserve(d.max_active);
    if (e != hipSuccess) {
        delete s;
        return hip_fail(e, "serve setup");
    }
    for (int i = (int)d.max_active - 1; i >= 0; i--) s->free_slots.push_back(i);
    s->t_start = now_s();
    c->served = true;  // direct calls on c now return VOSS_EBUSY
    s->th = std::thread(serve_loop, s);
    *out = s;
    return VOSS_OK;
}

int voss_serve_submit_interval(voss_server* s, uint64_t lo, uint64_t hi, uint64_t* id) {
    if (!s || !id) return fail(VOSS_EINVAL, "NULL argument");
    if (hi < lo || hi > (1ull << 62)) return fail(VOSS_EINVAL, "bad interval");
    auto r = std::make_shared<ServeReq>();
    r->lo = lo;
    r->hi = hi;
    r->t_submit = now_s();
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->err) return fail(s->err, s->errmsg);
    if (s->stopping) return fail(VOSS_EBUSY, "server is stopping");
    r->id = s->next_id++;
    s->queue.push_back(r);
    s->all[r->id] = r;
    *id = r->id;
    s->arrivals++;
    s->cv_work.notify_one();
    return VOSS_OK;
}

int voss_serve_submit_mr(voss_server* s, const uint64_t* n, size_t count, uint64_t* id) {
    if (!s || !id || (count && !n)) return fail(VOSS_EINVAL, "NULL argument");
    if (count >= 0xFFFFFFFFull) return fail(VOSS_EINVAL, "count must be below 2^32");
    auto r = std::make_shared<ServeReq>();
    r->is_mr = true;
    r->cand.assign(n, n + count);
    r->t_submit = now_s();
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->err) return fail(s->err, s->errmsg);
    if (s->stopping) return fail(VOSS_EBUSY, "server is stopping");
    r->id = s->next_id++;
    s->queue.push_back(r);
    s->all[r->id] = r;
    *id = r->id;
    s->arrivals++;
    s->cv_work.notify_one();
    return VOSS_OK;
}

int voss_serve_poll(voss_server* s, uint64_t id, int* done, voss_interval_result* out,
                    uint8_t* verdict, double* latency_s) {
    if (!s || !done) return fail(VOSS_EINVAL, "NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(id);
    if (it == s->all.end()) return fail(VOSS_ENOTFOUND, "unknown request id");
    *done = it->second->done ? 1 : 0;
    if (!*done) return VOSS_OK;
    return serve_result(s, it->second, out, verdict, latency_s);
}

int voss_serve_wait(voss_server* s, uint64_t id, voss_interval_result* out, uint8_t* verdict,
                    double* latency_s) {
    if (!s) return fail(VOSS_EINVAL, "NULL argument");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->all.find(id);
    if (it == s->all.end()) return fail(VOSS_ENOTFOUND, "unknown request id");
    ReqPtr r = it->second;
    s->cv_done.wait(lk, [&] { return r->done; });
    return serve_result(s, r, out, verdict, latency_s);
}

int voss_serve_stats(voss_server* s, uint64_t* steps, uint64_t* windows, double* device_busy_s,
                     double* wall_s) {
    if (!s) return fail(VOSS_EINVAL, "NULL argument");
    if (steps) *steps = s->n_steps.load();
    if (windows) *windows = s->n_windows.load();
    {
        std::lock_guard<std::mutex> g(s->stat_mu);
        if (device_busy_s) *device_busy_s = s->busy_ms / 1e3;
    }
    if (wall_s) *wall_s = now_s() - s->t_start;
    return VOSS_OK;
}

int voss_serve_stop(voss_server* s) {
    if (!s) return fail(VOSS_EINVAL, "NULL argument");
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stopping = true;
        s->cv_work.notify_all();
    }
    if (s->th.joinable()) s->th.join();
    const int err = s->err;
    Admit admit_(nullptr, true);
    s->c->served = false;
    (void)hipSetDevice(s->c->device);
    // the buffers stay with the context; nothing may be left in flight (after
    // a failure the scheduler can stop with launches it did not collect)
    (void)hipStreamSynchronize(s->c->stream);
    (void)hipStreamSynchronize(s->k->mr_stream);
    for (Step& S : s->k->steps) {
        S.X.reset();
        S.busy = false;
        S.by_slot.clear();
        S.finishing.clear();
    }
    for (MrLaunch& M : s->k->mr) {
        M.busy = false;
        M.reqs.clear();
    }
    delete s;
    return err;
}

}  // extern "C"
