// serve_steps.hip -- the host-stepped scheduler of the server
// (voss_serve_config.mode = VOSS_SERVE_HOST_STEPS): a sieve step is one
// executor plan of up to windows_per_step windows taken round-robin from
// every active interval request (bucket fill, sieve, stitch per round),
// followed on the same stream by a kernel that adds the step's sums into
// the requests' HBM slots.  Steps are enqueued without waiting for the
// device: there are two sets of step buffers, so while the device runs step
// k the host plans and enqueues step k+1, and the results of step k (pairs
// across step edges, large-gap candidates, overflow flags, finished
// requests) are collected when its event has completed.  Miller-Rabin
// batches are packed into one of two pinned staging sets, copied and tested
// on a second stream beside the sieve steps, and completed from their event.
// Results are exactly those of the request run alone: integer sums do not
// depend on order, and the pairs across step edges are joined on the host in
// window order.  A step whose bucket rooms or large-gap list overflowed
// restarts the requests it touched from their first window.
#include "serve_host.h"

namespace voss {
namespace serving {
namespace {

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
        finish_request(s, r.get());
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
void steps_loop_impl(voss_server* s) {
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
                finish_request(s, r.get());
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

}  // namespace

void steps_loop(voss_server* s) { steps_loop_impl(s); }

}  // namespace serving
}  // namespace voss
