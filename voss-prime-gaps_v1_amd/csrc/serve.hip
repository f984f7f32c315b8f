// serve.hip -- the server of the C ABI (voss_serve_*) and its default
// scheduler, the device-resident request loop.
//
// Requests (an interval, or a batch of Miller-Rabin candidates) wait in
// arrival order and are admitted while a request slot, the HBM budget and
// the loop's arenas allow.  An admitted interval request is planned once
// into LDS windows (with bucket tasks for its large primes), its records
// are written to pinned host memory, and its units (bucket tasks, then
// windows) are appended to the ring; an MR request's candidates are copied
// to pinned memory and its chunks appended.  The persistent loop kernel
// (serve_kernel.hip) claims units in ring order, so requests that arrive
// while others run are interleaved with them on the device as soon as they
// are admitted; a request's sums stay in its HBM slot until the workgroup
// that completes its last unit writes them to the request's mailbox.  The
// scheduler thread here only admits, publishes and collects: it never waits
// for the device while requests run, and relaunches the loop when it has
// exited for lack of work.
//
// Results are exactly those of the request run alone (integer sums do not
// depend on order; the pairs across windows are joined on the device in
// window order).  A request whose bucket rooms or large-gap list overflowed
// starts over with twice the rooms or a larger list.
#include <cstring>

#include "serve_host.h"

namespace voss {
namespace serving {

hipError_t PinnedBuf::reserve(size_t n) {
    if (n <= bytes && p) return hipSuccess;
    release();
    hipError_t e = hipHostMalloc(&p, n, hipHostMallocCoherent | hipHostMallocMapped);
    if (e != hipSuccess) {
        p = nullptr;
        return e;
    }
    bytes = n;
    std::memset(p, 0, n);
    return hipSuccess;
}

void PinnedBuf::release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    bytes = 0;
}

void LoopBufs::release() {
    ring.release(); ctl.release(); req.release(); mail.release(); mrq.release();
    mr_mail.release(); host.release(); claimed.release(); wgstat.release();
    d_ctr.release(); d_slot.release(); d_mr_done.release(); d_acc.release(); d_ent.release();
    d_rec.release(); d_cur.release(); d_tile.release();
    if (k_start) (void)hipEventDestroy(k_start);
    if (k_done) (void)hipEventDestroy(k_done);
    k_start = k_done = nullptr;
    ready = false;
}

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
    k->loop.release();
    delete k;
}

uint64_t req_hbm(const voss_server* s, const ServeReq& r) {
    if (r.is_mr) return r.cand.size() * (8 + 1 + 4);
    return sizeof(JobAcc) + (uint64_t)s->cfg.large_capacity * sizeof(LargeRaw);
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

void finish_request(voss_server* s, ServeReq* r) {
    r->large.resize(s->cfg.large_capacity);
    r->res.large = r->large.data();
    r->res.large_capacity = s->cfg.large_capacity;
    finish_interval(r->J, r->lo, r->hi, &r->res);
    r->large.resize(std::min<uint64_t>(r->res.large_count, s->cfg.large_capacity));
    if (r->J.overflow) r->status = VOSS_ERANGE;
}

// ------------------------------------------------------- device-resident loop

namespace {

constexpr uint64_t kRingCap = 1ull << 20;       // units
constexpr uint32_t kMrSlots = 64;
constexpr uint64_t kHostArenaMin = 256ull << 20;  // pinned bytes
constexpr uint64_t kEntriesMax = (1ull << 32) - 64;  // bucket offsets are 32-bit
constexpr uint64_t kRecMin = 1ull << 20;         // windows
constexpr uint64_t kTileMin = 1ull << 14;
constexpr uint32_t kIdleLimit = 16384;           // empty polls (a few us each) before exit
constexpr uint32_t kDepLimit = 1u << 22;         // polls of a bucket dependency (seconds)
constexpr uint32_t kLoopWindow = 131072;         // the loop kernel's window
constexpr uint32_t kAheadPerWorkgroup = 1;       // units published ahead of the device
constexpr uint32_t kBurst = 8;                   // units of one request per turn
// A tile is split into bucket tasks of at least this estimated cost
// (bucket_task_cost units): ~0.35 ms of a workgroup when the device has
// little queued work, so one request spreads over many workgroups, and
// ~1.4 ms when units are piling up (a split tile's tasks count their hits
// before they write them, so fewer, larger tasks do less work: measured on
// a Poisson stream of intervals in [1e10, 1e16) of lengths 1e5..1e10 mixed
// with Miller-Rabin batches, the loop's bucket work fell from 20 to 11
// workgroup-seconds between the two).
constexpr double kMinTaskCostIdle = 1.5e7;
constexpr double kMinTaskCostBusy = 6e7;
constexpr uint32_t kMaxBypass = 256;             // admissions that may pass a waiting request
constexpr uint64_t kEarlyWindows = 4;            // windows of a request that may wait on the device

template <typename T>
T* at(PinnedBuf& b, uint64_t byte_off) {
    return reinterpret_cast<T*>(static_cast<char*>(b.p) + byte_off);
}

struct Loop {
    voss_server* s;
    voss_ctx* c;
    LoopBufs& L;
    RangeAlloc ent, rec, cur, tile, host;
    std::vector<ReqPtr> live;  // admitted, not yet complete
    std::vector<ReqPtr> feed;  // admitted, with units not yet published
    std::vector<int> mr_free;
    std::vector<uint32_t> slot_gen, mr_gen;
    uint64_t tail = 0;
    bool kernel_live = false;
    uint64_t dev_used = 0;     // device bytes of the live interval requests
    Loop(voss_server* sv, voss_ctx* cx, LoopBufs& lb) : s(sv), c(cx), L(lb) {}

    ServeCtl* ctl() { return static_cast<ServeCtl*>(L.ctl.p); }

    // Waits for the loop kernel to exit (stop = 2: at its next claim).
    int quiesce() {
        if (!kernel_live) return VOSS_OK;
        __atomic_store_n(&ctl()->stop, 2u, __ATOMIC_RELEASE);
        HIPCHK(hipEventSynchronize(L.k_done));
        __atomic_store_n(&ctl()->stop, 0u, __ATOMIC_RELEASE);
        return reap();
    }

    // The kernel has exited: its time into the statistics.
    int reap() {
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, L.k_start, L.k_done));
        {
            std::lock_guard<std::mutex> g(s->stat_mu);
            s->busy_ms += ms;
        }
        kernel_live = false;
        return VOSS_OK;
    }

    // (Re)allocates the arenas (no kernel running, nothing live) and
    // brings every counter to its free state.
    int setup(uint64_t ent_n, uint64_t rec_n, uint64_t tile_n, uint64_t host_b) {
        const uint32_t A = s->cfg.max_active;
        L.max_active = A;
        L.mr_slots = kMrSlots;
        L.ring_cap = kRingCap;
        HIPCHK(L.ring.reserve(kRingCap * sizeof(ServeUnit)));
        HIPCHK(L.ctl.reserve(sizeof(ServeCtl)));
        HIPCHK(L.req.reserve(A * sizeof(ServeReqInfo)));
        HIPCHK(L.mail.reserve(A * sizeof(ServeMailbox)));
        HIPCHK(L.mrq.reserve(kMrSlots * sizeof(ServeMrInfo)));
        HIPCHK(L.mr_mail.reserve(kMrSlots * sizeof(unsigned long long)));
        HIPCHK(L.host.reserve(host_b));
        HIPCHK(L.claimed.reserve(std::max<size_t>(1024, (size_t)c->num_cus) * 8));
        std::memset(L.claimed.p, 0, L.claimed.bytes);
        HIPCHK(L.wgstat.reserve(std::max<size_t>(1024, (size_t)c->num_cus) * 64));
        std::memset(L.wgstat.p, 0, L.wgstat.bytes);
        int khz = 0;
        if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, c->device) == hipSuccess &&
            khz > 0)
            L.clock_hz = khz * 1e3;
        std::memset(L.mail.p, 0, L.mail.bytes);
        std::memset(L.mr_mail.p, 0, L.mr_mail.bytes);
        std::memset(L.ctl.p, 0, L.ctl.bytes);
        HIPCHK(L.d_ctr.reserve(64));  // claim counters on separate lines, units done
        HIPCHK(L.d_slot.reserve(A));
        HIPCHK(L.d_mr_done.reserve(kMrSlots));
        HIPCHK(L.d_acc.reserve(A));
        HIPCHK(L.d_ent.reserve(ent_n));
        HIPCHK(L.d_rec.reserve(rec_n));
        HIPCHK(L.d_cur.reserve(rec_n));
        HIPCHK(L.d_tile.reserve(tile_n));
        if (!L.k_start) HIPCHK(hipEventCreate(&L.k_start));
        if (!L.k_done) HIPCHK(hipEventCreate(&L.k_done));
        hipStream_t st = c->stream;
        HIPCHK(launch_fill_u32(reinterpret_cast<uint32_t*>(L.d_ctr.p), 128, 0, st));
        HIPCHK(launch_fill_u32(reinterpret_cast<uint32_t*>(L.d_slot.p),
                               L.d_slot.cap * sizeof(ServeSlot) / 4, 0, st));
        HIPCHK(launch_fill_u32(L.d_mr_done.p, L.d_mr_done.cap, 0, st));
        HIPCHK(launch_fill_u32(L.d_cur.p, L.d_cur.cap, 0, st));
        HIPCHK(launch_fill_u32(L.d_tile.p, L.d_tile.cap, 0, st));
        HIPCHK(launch_init_acc(L.d_acc.p, (uint32_t)L.d_acc.cap, st));
        HIPCHK(hipStreamSynchronize(st));
        ent.reset(std::min<uint64_t>(L.d_ent.cap, kEntriesMax));
        rec.reset(L.d_rec.cap);
        cur.reset(L.d_cur.cap);
        tile.reset(L.d_tile.cap);
        host.reset(L.host.bytes);
        tail = 0;
        L.ready = true;
        return VOSS_OK;
    }

    int launch() {
        SieveArgs a = make_args(c, c->eb, 0, 0);
        ServeArgs sv{};
        sv.ring = static_cast<const ServeUnit*>(L.ring.p);
        sv.ring_mask = L.ring_cap - 1;
        sv.ctl = ctl();
        sv.claim = L.d_ctr.p;
        sv.req = static_cast<const ServeReqInfo*>(L.req.p);
        sv.mrq = static_cast<const ServeMrInfo*>(L.mrq.p);
        sv.mail = static_cast<ServeMailbox*>(L.mail.p);
        sv.mr_mail = static_cast<unsigned long long*>(L.mr_mail.p);
        sv.slot = L.d_slot.p;
        sv.mr_done = L.d_mr_done.p;
        sv.acc = L.d_acc.p;
        sv.entries = L.d_ent.p;
        sv.units_done = L.d_ctr.p + 48;
        sv.claimed_by = static_cast<unsigned long long*>(L.claimed.p);
        sv.wg_stats = static_cast<unsigned long long*>(L.wgstat.p);
        sv.idle_limit = kIdleLimit;
        sv.dep_limit = kDepLimit;
        sv.log2w = log2_window(c);
        // the device's copy of the stop flag and the poll ticket start clear
        // (an earlier launch may have stopped on them)
        HIPCHK(launch_fill_u32(reinterpret_cast<uint32_t*>(L.d_ctr.p + 17), 2, 0, c->stream));
        HIPCHK(launch_fill_u32(reinterpret_cast<uint32_t*>(L.d_ctr.p + 32), 2, 0, c->stream));
        HIPCHK(hipEventRecord(L.k_start, c->stream));
        HIPCHK(launch_serve(a, sv, (uint32_t)c->num_cus, c->stream));
        HIPCHK(hipEventRecord(L.k_done, c->stream));
        kernel_live = true;
        s->n_steps++;
        return VOSS_OK;
    }

    // A lower bound of the units the device has claimed.
    uint64_t claimed() const {
        const unsigned long long* cb = static_cast<const unsigned long long*>(L.claimed.p);
        uint64_t m = 0;
        for (int i = 0; i < c->num_cus; i++) m = std::max<uint64_t>(m, __atomic_load_n(&cb[i], __ATOMIC_ACQUIRE));
        return m;
    }

    // Units of the admitted requests that the device has not claimed yet
    // (published ahead, or still to be published).
    uint64_t queued_units() const {
        uint64_t n = tail - claimed();
        for (const auto& r : feed) n += (r->units.size() - r->upub) + (r->nwin - r->wpub);
        return n;
    }

    // The next unit request r may publish now: a window whose tile's bucket
    // tasks are done (or that has none, or, for an early request, are all
    // published), else its next bucket task or MR chunk.  false when it has
    // none ready.
    bool next_unit(ServeReq* r, ServeUnit* u) {
        if (!r->is_mr && r->wpub < r->nwin) {
            const uint32_t t = r->win_tile[r->wpub];
            if (t == ~0u || (r->early && r->upub >= r->tile_end[t]) ||
                __atomic_load_n(&r->tile_ready[t], __ATOMIC_ACQUIRE)) {
                *u = {kUnitWindow, (uint32_t)r->slot, (uint32_t)r->wpub++, r->gen, 0};
                return true;
            }
        }
        if (r->upub < r->units.size()) {
            *u = r->units[r->upub++];
            return true;
        }
        return false;
    }

    // Keeps about kAheadPerWorkgroup units per workgroup published ahead of
    // the device.  Each burst of up to kBurst units goes to the request that
    // has had the fewest units published so far among those with a unit
    // ready (least attained service: a short request that arrives behind a
    // long one is fed first).  Measured on that mixed stream at 1.3x the
    // plain path's capacity: p50 1.16 -> 0.92 ms and p99 16.9 -> 13.1 ms
    // against two units ahead in plain round-robin, at the same rate.
    void publish_some() {
        if (feed.empty()) return;
        const uint64_t ahead = (uint64_t)kAheadPerWorkgroup * c->num_cus;
        const uint64_t cl = claimed();
        ServeUnit* ring = static_cast<ServeUnit*>(L.ring.p);
        const uint64_t t0 = tail;
        std::vector<uint8_t> blocked(feed.size(), 0);  // nothing ready in this call
        while (tail - cl < ahead) {
            size_t k = feed.size();
            for (size_t j = 0; j < feed.size(); j++) {
                if (blocked[j]) continue;
                if (k == feed.size() || feed[j]->upub + feed[j]->wpub < feed[k]->upub + feed[k]->wpub) k = j;
            }
            if (k == feed.size()) break;
            ServeReq* r = feed[k].get();
            uint32_t n = 0;
            ServeUnit u;
            while (n < kBurst && tail - cl < ahead && next_unit(r, &u)) {
                u.seq = tail;
                ring[(tail++) & (L.ring_cap - 1)] = u;
                n++;
            }
            if (r->upub == r->units.size() && (r->is_mr || r->wpub == r->nwin)) {
                feed.erase(feed.begin() + (long)k);
                blocked.erase(blocked.begin() + (long)k);
            } else if (n < kBurst && tail - cl < ahead) {
                blocked[k] = 1;
            }
        }
        if (tail != t0) __atomic_store_n(&ctl()->tail, (unsigned long long)tail, __ATOMIC_RELEASE);
    }

    void free_ranges(ServeReq* r) {
        ent.release(r->ent_off, r->ent_n);
        rec.release(r->rec_off, r->rec_n);
        cur.release(r->cur_off, r->cur_n);
        tile.release(r->tile_off, r->tile_n);
        host.release(r->host_off, r->host_n);
        r->ent_n = r->rec_n = r->cur_n = r->tile_n = r->host_n = 0;
    }

    // Plans an interval request and queues its units for publication.
    // *placed = false: it does not fit now (arenas, budget) and stays queued.
    int admit_interval(const ReqPtr& rp, bool alone, bool* placed) {
        ServeReq* r = rp.get();
        *placed = false;
        const uint32_t W = kLoopWindow;
        if (r->win.empty() && r->hi > 7) {
            const uint64_t top = r->hi / 30 * 30 + 30ull * W + 30;
            const uint64_t need = isqrt_u64(top) + 1;
            if (need > c->primes_limit || c->d_nprimes != c->primes.size()) {
                // the base-prime table moves: only with nothing on the device
                if (!alone) return VOSS_OK;
                int rc = quiesce();
                if (rc || (rc = ensure_primes(c, top))) return rc;
            }
            Plan P1;
            P1.njobs = 1;
            const uint64_t a = std::max<uint64_t>(r->lo, 7), b = r->hi - 1;
            plan_range(c, P1, a, b, (uint32_t)r->slot, 0, W, (a / 30) & ~4095ull);
            r->win.swap(P1.desc);
        }
        if (r->win.empty()) {
            finish_request(s, r);
            *placed = true;
            return VOSS_OK;
        }
        // bucket tiles and rooms (one round, or no buckets for a request
        // beyond a round)
        Plan P;
        P.njobs = s->cfg.max_active;
        P.desc = r->win;
        std::vector<Round> rounds;
        const uint32_t pct0 = c->tune.bucket_room_pct;
        if (r->room_pct) c->tune.bucket_room_pct = r->room_pct;
        make_rounds(c, P, rounds,
                    queued_units() > 4ull * (uint64_t)c->num_cus ? kMinTaskCostBusy : kMinTaskCostIdle);
        c->tune.bucket_room_pct = pct0;
        const bool bucketed = rounds.size() == 1 && !rounds[0].desc.empty() &&
                              rounds[0].room_total <= kEntriesMax / 2;
        const DescVec& desc = bucketed ? rounds[0].desc : r->win;
        if (!bucketed)
            for (const SegDesc& d : desc)
                if (d.bucket != kNoBucket) return fail(VOSS_EDEVICE, "serve: unexpected bucket");
        const Round* R = bucketed ? &rounds[0] : nullptr;
        const uint64_t nwin = desc.size();
        const uint64_t nbkt = R ? R->nbuckets : 0;
        const uint64_t ntask = R ? R->tasks.size() : 0;
        const uint64_t ntile = (nbkt + kBucketTileMax - 1) / kBucketTileMax;
        const uint64_t nent = R ? R->room_total : 0;
        if (!r->lcap) {
            const double n = (double)(r->hi - r->lo);
            const double primes = n / std::max(1.0, std::log((double)r->hi + 3.0));
            r->lcap = std::max<uint64_t>(1024, (uint64_t)(4e-5 * primes));
        }
        const uint64_t units = ntask + nwin;
        const uint64_t dev_b = nent * 4 + nwin * (sizeof(SegRec) + 4) + ntile * 4 + sizeof(JobAcc);
        auto up16 = [](uint64_t x) { return (x + 15) / 16 * 16; };
        const uint64_t hb_desc = 0, hb_task = hb_desc + up16(nwin * sizeof(SegDesc));
        const uint64_t hb_boff = hb_task + up16(ntask * sizeof(BucketTask));
        const uint64_t hb_tnt = hb_boff + up16((nbkt + 1) * 8);
        const uint64_t hb_ready = hb_tnt + up16(ntile * 4);
        const uint64_t hb_large = hb_ready + up16(ntile * 4);
        const uint64_t hbytes = hb_large + up16(r->lcap * sizeof(LargeRaw));
        if (!alone && dev_used + dev_b > s->cfg.hbm_budget_bytes) return VOSS_OK;
        bool ok = ent.alloc(nent, 4, &r->ent_off);
        if (ok) r->ent_n = nent;
        ok = ok && rec.alloc(nwin, 1, &r->rec_off);
        if (ok) r->rec_n = nwin;
        ok = ok && cur.alloc(nbkt, 1, &r->cur_off);
        if (ok) r->cur_n = nbkt;
        ok = ok && tile.alloc(ntile, 1, &r->tile_off);
        if (ok) r->tile_n = ntile;
        ok = ok && host.alloc(hbytes, 256, &r->host_off);
        if (ok) r->host_n = hbytes;
        if (!ok) {
            free_ranges(r);
            if (!alone) return VOSS_OK;
            // nothing live: grow the arenas to hold it and try again
            int rc = quiesce();
            if (rc) return rc;
            const uint64_t e2 = std::min<uint64_t>(std::max<uint64_t>(L.d_ent.cap, nent + 64), kEntriesMax);
            const uint64_t r2 = std::max<uint64_t>(L.d_rec.cap, nwin + nbkt + 64);
            const uint64_t t2 = std::max<uint64_t>(L.d_tile.cap, ntile + 64);
            const uint64_t h2 = std::max<uint64_t>(L.host.bytes, hbytes + 4096);
            if ((rc = setup(e2, r2, t2, h2))) return rc;
            if (!ent.alloc(nent, 4, &r->ent_off) || !rec.alloc(nwin, 1, &r->rec_off) ||
                !cur.alloc(nbkt, 1, &r->cur_off) || !tile.alloc(ntile, 1, &r->tile_off) ||
                !host.alloc(hbytes, 256, &r->host_off))
                return fail(VOSS_ENOMEM, "serve: request beyond the arenas");
            r->ent_n = nent;
            r->rec_n = nwin;
            r->cur_n = nbkt;
            r->tile_n = ntile;
            r->host_n = hbytes;
        }
        // the request's host records
        const uint64_t hb = r->host_off;
        SegDesc* hdesc = at<SegDesc>(L.host, hb + hb_desc);
        std::memcpy(hdesc, desc.data(), nwin * sizeof(SegDesc));
        BucketTask* htask = at<BucketTask>(L.host, hb + hb_task);
        unsigned long long* hboff = at<unsigned long long>(L.host, hb + hb_boff);
        uint32_t* htnt = at<uint32_t>(L.host, hb + hb_tnt);
        uint32_t* hready = at<uint32_t>(L.host, hb + hb_ready);
        for (uint64_t t = 0; t < ntile; t++) hready[t] = 0;
        if (R) {
            std::memcpy(htask, R->tasks.data(), ntask * sizeof(BucketTask));
            unsigned long long o = r->ent_off;
            for (uint64_t b = 0; b < nbkt; b++) {
                hboff[b] = o;
                o += R->room[b];
            }
            hboff[nbkt] = o;
            for (uint64_t t = 0; t < ntile; t++) htnt[t] = 0;
            for (const BucketTask& T : R->tasks) htnt[T.b0 / kBucketTileMax]++;
        }
        r->gen = ++slot_gen[r->slot];
        ServeReqInfo& I = static_cast<ServeReqInfo*>(L.req.p)[r->slot];
        I.desc = hdesc;
        I.tasks = htask;
        I.boff = hboff;
        I.tile_ntask = htnt;
        I.tile_ready = hready;
        I.cursor = L.d_cur.p + r->cur_off;
        I.tile_done = L.d_tile.p + r->tile_off;
        I.rec = L.d_rec.p + r->rec_off;
        I.large = at<LargeRaw>(L.host, hb + hb_large);
        I.lcap = r->lcap;
        I.nwin = (uint32_t)nwin;
        I.ntask = (uint32_t)ntask;
        I.nbkt = (uint32_t)nbkt;
        I.ntile = (uint32_t)ntile;
        I.gen = r->gen;
        // units: the bucket tasks in tile order; the windows follow their
        // tile's tasks once the device has finished them (publish_some), or,
        // for a request of at most kEarlyWindows windows admitted onto an
        // idle device, as soon as the tasks are published: the window then
        // waits for them on the device, and the request skips a round trip
        // through the host.  Measured on an idle server (median of 25): a
        // 1e5-number interval at 1e13 0.37 -> 0.23 ms; for 26 windows
        // (1e8 numbers at 1e15) the waiting windows made it slower (0.67 ->
        // 0.89 ms), and under load they held workgroups that other units
        // needed, so larger requests keep the host gate
        r->units.clear();
        r->upub = 0;
        r->tile_end.assign(ntile, 0);
        r->early = nwin <= kEarlyWindows && queued_units() < (uint64_t)c->num_cus / 4;
        if (R) {
            std::vector<std::vector<uint32_t>> tasks_of(ntile);
            for (uint64_t k = 0; k < ntask; k++)
                tasks_of[R->tasks[k].b0 / kBucketTileMax].push_back((uint32_t)k);
            r->units.reserve(ntask);
            for (uint64_t t = 0; t < ntile; t++) {
                for (uint32_t k : tasks_of[t])
                    r->units.push_back({kUnitBucket, (uint32_t)r->slot, k, r->gen, 0});
                r->tile_end[t] = (uint32_t)r->units.size();
            }
        }
        r->nwin = nwin;
        r->wpub = 0;
        r->win_tile.resize(nwin);
        for (uint64_t w = 0; w < nwin; w++)
            r->win_tile[w] = desc[w].bucket == kNoBucket ? ~0u : desc[w].bucket / kBucketTileMax;
        r->tile_ready = hready;
        r->dev_bytes = dev_b;
        dev_used += dev_b;
        s->n_windows += nwin;
        live.push_back(rp);
        feed.push_back(rp);
        *placed = true;
        return VOSS_OK;
    }

    int admit_mr(const ReqPtr& rp, bool alone, bool* placed) {
        ServeReq* r = rp.get();
        *placed = false;
        const uint64_t n = r->cand.size();
        if (n == 0) {
            *placed = true;
            return VOSS_OK;
        }
        if (mr_free.empty()) return VOSS_OK;
        const uint64_t nchunk = (n + kServeMrChunk - 1) / kServeMrChunk;
        const uint64_t hbytes = n * 8 + (n + 3) / 4 * 4;
        if (!host.alloc(hbytes, 256, &r->host_off)) {
            if (!alone) return VOSS_OK;
            int rc = quiesce();
            if (rc) return rc;
            if ((rc = setup(L.d_ent.cap, L.d_rec.cap, L.d_tile.cap,
                            std::max<uint64_t>(L.host.bytes, hbytes + 4096))))
                return rc;
            if (!host.alloc(hbytes, 256, &r->host_off)) return fail(VOSS_ENOMEM, "serve: MR batch beyond the arena");
        }
        r->host_n = hbytes;
        r->slot = mr_free.back();
        mr_free.pop_back();
        uint64_t* hc = at<uint64_t>(L.host, r->host_off);
        std::memcpy(hc, r->cand.data(), n * 8);
        r->gen = ++mr_gen[r->slot];
        ServeMrInfo& M = static_cast<ServeMrInfo*>(L.mrq.p)[r->slot];
        M.cand = hc;
        M.verdict = at<uint8_t>(L.host, r->host_off + n * 8);
        M.count = n;
        M.nchunk = (uint32_t)nchunk;
        M.gen = r->gen;
        r->units.clear();
        r->upub = 0;
        for (uint64_t k = 0; k < nchunk; k++)
            r->units.push_back({kUnitMr, (uint32_t)r->slot, (uint32_t)k, r->gen, 0});
        live.push_back(rp);
        feed.push_back(rp);
        *placed = true;
        return VOSS_OK;
    }

    // Collects the requests whose mailbox shows their generation.  restart
    // receives the ones that must run again (overflow).
    int collect(std::vector<ReqPtr>& finished, std::vector<ReqPtr>& restart) {
        for (size_t i = 0; i < live.size();) {
            const ReqPtr rp = live[i];
            ServeReq* r = rp.get();
            bool done;
            if (r->is_mr) {
                unsigned long long* mm = static_cast<unsigned long long*>(L.mr_mail.p);
                done = __atomic_load_n(&mm[r->slot], __ATOMIC_ACQUIRE) == r->gen;
            } else {
                ServeMailbox* mb = static_cast<ServeMailbox*>(L.mail.p) + r->slot;
                done = __atomic_load_n(&mb->done_gen, __ATOMIC_ACQUIRE) == r->gen;
            }
            if (!done) {
                i++;
                continue;
            }
            live[i] = live.back();
            live.pop_back();
            if (r->is_mr) {
                const uint64_t n = r->cand.size();
                const uint8_t* v = at<uint8_t>(L.host, r->host_off + n * 8);
                r->verdict.assign(v, v + n);
                host.release(r->host_off, r->host_n);
                r->host_n = 0;
                mr_free.push_back(r->slot);
                finished.push_back(rp);
                continue;
            }
            const ServeMailbox* mb = static_cast<const ServeMailbox*>(L.mail.p) + r->slot;
            dev_used -= r->dev_bytes;
            if (mb->error) {
                free_ranges(r);
                return fail(VOSS_EDEVICE, "serve: a window's bucket dependency did not complete");
            }
            if (mb->overflow || mb->large_n > r->lcap) {
                if (mb->overflow) r->room_pct = 2 * (r->room_pct ? r->room_pct : c->tune.bucket_room_pct);
                if (mb->large_n > r->lcap) r->lcap = mb->large_n + mb->large_n / 4 + 64;
                free_ranges(r);
                restart.push_back(rp);
                continue;
            }
            r->J = JobHost();
            add_acc(r->J, mb->acc);
            if (mb->acc.error) r->J.overflow = true;
            const ServeReqInfo& I = static_cast<const ServeReqInfo*>(L.req.p)[r->slot];
            const LargeRaw* raw = I.large;
            r->J.large.resize(mb->large_n);
            for (uint64_t k = 0; k < mb->large_n; k++)
                r->J.large[k] = {raw[k].p_before + raw[k].gap, (int32_t)raw[k].gap,
                                 merit_of(raw[k].p_before, raw[k].gap)};
            free_ranges(r);
            finish_request(s, r);
            finished.push_back(rp);
        }
        return VOSS_OK;
    }
};

}  // namespace

void device_loop(voss_server* s) {
    voss_ctx* c = s->c;
    t_scheduler = true;
    if (hipSetDevice(c->device) != hipSuccess) {
        serve_fail_all(s, VOSS_EDEVICE, "hipSetDevice");
        return;
    }
    Loop lp(s, c, s->k->loop);
    lp.slot_gen.assign(s->cfg.max_active, 0);
    lp.mr_gen.assign(kMrSlots, 0);
    for (int i = (int)kMrSlots - 1; i >= 0; i--) lp.mr_free.push_back(i);
    int rc = VOSS_OK;
    {
        gate().enter();
        const uint64_t budget = s->cfg.hbm_budget_bytes;
        // bucket entries take most of the budget (4 bytes each)
        const uint64_t ent_n = std::min<uint64_t>(std::max<uint64_t>(budget / 20 * 3, 1ull << 24), kEntriesMax);
        rc = lp.setup(std::max<uint64_t>(ent_n, s->k->loop.d_ent.cap),
                      std::max<uint64_t>(kRecMin, s->k->loop.d_rec.cap),
                      std::max<uint64_t>(kTileMin, s->k->loop.d_tile.cap),
                      std::max<uint64_t>(kHostArenaMin, s->k->loop.host.bytes));
        gate().leave();
    }
    // requests taken from the queue and not yet published (arrival order)
    std::deque<ReqPtr> pending;
    bool idle_poll = false;
    while (rc == VOSS_OK) {
        if (idle_poll) {
            // requests on the device: poll their mailboxes again shortly
            // (new arrivals wake the scheduler at once)
            std::unique_lock<std::mutex> lk(s->mu);
            const uint64_t seen = s->arrivals.load();
            s->cv_work.wait_for(lk, std::chrono::microseconds(20),
                                [&] { return s->arrivals.load() != seen; });
        }
        {
            std::unique_lock<std::mutex> lk(s->mu);
            if (lp.live.empty() && pending.empty() && s->queue.empty()) {
                if (s->stopping) break;
                // nothing to run: wait for a request (the loop kernel exits
                // by itself once it finds the ring empty)
                s->cv_work.wait_for(lk, std::chrono::milliseconds(50), [&] {
                    return s->stopping || !s->queue.empty();
                });
            }
            // requests into slots while slots are free, in arrival order
            while (!s->queue.empty()) {
                ReqPtr r = s->queue.front();
                if (!r->is_mr) {
                    if (s->free_slots.empty()) break;
                    r->slot = s->free_slots.back();
                    s->free_slots.pop_back();
                }
                r->hbm = 0;
                pending.push_back(r);
                s->queue.pop_front();
            }
        }
        {
            struct PassGate {
                PassGate() { gate().enter(); }
                ~PassGate() { gate().leave(); }
            } pass_gate;
            // admit pending requests in arrival order; one that does not fit
            // (its bucket entries, the budget) is passed by later ones that
            // do, at most kMaxBypass times, after which nothing passes it
            std::vector<ReqPtr> immediate;
            for (auto it = pending.begin(); rc == VOSS_OK && it != pending.end();) {
                ReqPtr r = *it;
                const bool alone = lp.live.empty();
                bool placed = false;
                rc = r->is_mr ? lp.admit_mr(r, alone, &placed) : lp.admit_interval(r, alone, &placed);
                if (rc) break;
                if (!placed) {
                    if (r->bypassed >= kMaxBypass) break;
                    r->bypassed++;
                    ++it;
                    continue;
                }
                it = pending.erase(it);
                if ((r->is_mr && r->cand.empty()) || (!r->is_mr && r->win.empty())) immediate.push_back(r);
            }
            if (rc) break;
            complete(s, immediate);
            lp.publish_some();
            // the loop kernel runs while there is work for it
            if (lp.kernel_live) {
                const hipError_t q = hipEventQuery(s->k->loop.k_done);
                if (q == hipSuccess) {
                    if ((rc = lp.reap())) break;
                } else if (q != hipErrorNotReady) {
                    rc = hip_fail(q, "serve loop kernel");
                    break;
                }
            }
            if (!lp.live.empty() && !lp.kernel_live && (rc = lp.launch())) break;
            if (__atomic_load_n(&lp.ctl()->fault, __ATOMIC_ACQUIRE)) {
                rc = fail(VOSS_EDEVICE, "serve: a ring entry was claimed before it was written");
                break;
            }
            // results
            std::vector<ReqPtr> finished, again;
            {
                std::lock_guard<std::mutex> lk(s->mu);
                rc = lp.collect(finished, again);
            }
            if (rc) break;
            for (auto it = again.rbegin(); it != again.rend(); ++it) pending.push_front(*it);
            complete(s, finished);
            lp.publish_some();
            idle_poll = !lp.live.empty() && finished.empty() && again.empty();
        }
    }
    {
        gate().enter();
        if (rc == VOSS_OK && lp.kernel_live) {
            __atomic_store_n(&lp.ctl()->stop, 1u, __ATOMIC_RELEASE);
            const hipError_t e = hipEventSynchronize(s->k->loop.k_done);
            __atomic_store_n(&lp.ctl()->stop, 0u, __ATOMIC_RELEASE);
            if (e != hipSuccess) rc = hip_fail(e, "serve loop kernel");
            else rc = lp.reap();
        }
        if (rc != VOSS_OK) {
            if (lp.kernel_live) {
                __atomic_store_n(&lp.ctl()->stop, 2u, __ATOMIC_RELEASE);
                (void)hipEventSynchronize(s->k->loop.k_done);
                __atomic_store_n(&lp.ctl()->stop, 0u, __ATOMIC_RELEASE);
                lp.kernel_live = false;
            }
            // the arenas' state is unknown after a failure: set up again next time
            s->k->loop.ready = false;
        }
        gate().leave();
    }
    if (rc != VOSS_OK) serve_fail_all(s, rc, voss_last_error());
}

namespace {

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
}  // namespace serving
}  // namespace voss

using namespace voss::serving;

extern "C" {

int voss_serve_start(voss_ctx* c, const voss_serve_config* cfg, voss_server** out) {
    if (!c || !out) return fail(VOSS_EINVAL, "NULL argument");
    *out = nullptr;
    VOSS_ADMIT(c);  // also refuses a context that already has a server
    int rc = set_device(c);
    if (rc) return rc;
    voss_serve_config d{64, 4096, 32ull << 30, 64, 1u << 22, VOSS_SERVE_DEVICE_LOOP};
    if (cfg) {
        if (cfg->max_active) d.max_active = cfg->max_active;
        if (cfg->windows_per_step) d.windows_per_step = cfg->windows_per_step;
        if (cfg->hbm_budget_bytes) d.hbm_budget_bytes = cfg->hbm_budget_bytes;
        if (cfg->large_capacity) d.large_capacity = cfg->large_capacity;
        if (cfg->mr_per_step) d.mr_per_step = cfg->mr_per_step;
        d.mode = cfg->mode;
    }
    if (d.mode != VOSS_SERVE_DEVICE_LOOP && d.mode != VOSS_SERVE_HOST_STEPS)
        return fail(VOSS_EINVAL, "unknown serving mode");
    voss_server* s = new voss_server();
    s->c = c;
    s->cfg = d;
    if (d.mode == VOSS_SERVE_DEVICE_LOOP &&
        (c->tune.window_bytes != kLoopWindow || c->tune.wg_per_cu != 1)) {
        s->saved_tune = c->tune;
        s->tune_saved = true;
        c->tune.window_bytes = kLoopWindow;
        c->tune.wg_per_cu = 1;
    }
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
            if (s->tune_saved) c->tune = s->saved_tune;
            delete s;
            return hip_fail(e, "serve setup");
        }
        c->serve_cache = k;
        c->free_serve_cache = free_serve_cache;
    }
    s->k = static_cast<ServeCache*>(c->serve_cache);
    if (d.mode == VOSS_SERVE_HOST_STEPS) {
        const hipError_t e = s->k->d_acc.reserve(d.max_active);
        if (e != hipSuccess) {
            if (s->tune_saved) c->tune = s->saved_tune;
            delete s;
            return hip_fail(e, "serve setup");
        }
    }
    for (int i = (int)d.max_active - 1; i >= 0; i--) s->free_slots.push_back(i);
    s->t_start = now_s();
    c->served = true;  // direct calls on c now return VOSS_EBUSY
    s->th = std::thread(d.mode == VOSS_SERVE_HOST_STEPS ? steps_loop : device_loop, s);
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

int voss_serve_unit_stats(voss_server* s, uint64_t* units, double* busy_s) {
    if (!s || !units || !busy_s) return fail(VOSS_EINVAL, "NULL argument");
    for (int k = 0; k < 3; k++) {
        units[k] = 0;
        busy_s[k] = 0;
    }
    if (s->cfg.mode != VOSS_SERVE_DEVICE_LOOP) return VOSS_OK;
    // per workgroup and kind: units run, wall-clock ticks (the device writes
    // them after each unit)
    const LoopBufs& L = s->k->loop;
    const unsigned long long* w = static_cast<const unsigned long long*>(L.wgstat.p);
    for (int g = 0; g < s->c->num_cus; g++)
        for (int k = 0; k < 3; k++) {
            units[k] += __atomic_load_n(&w[8 * g + 2 * k], __ATOMIC_RELAXED);
            busy_s[k] += (double)__atomic_load_n(&w[8 * g + 2 * k + 1], __ATOMIC_RELAXED) / L.clock_hz;
        }
    return VOSS_OK;
}

int voss_serve_stats(voss_server* s, uint64_t* steps, uint64_t* windows, double* device_busy_s,
                     double* wall_s) {
    if (!s) return fail(VOSS_EINVAL, "NULL argument");
    if (steps) *steps = s->n_steps.load();
    if (windows) *windows = s->n_windows.load();
    const double now = now_s();
    if (device_busy_s) {
        if (s->cfg.mode == VOSS_SERVE_DEVICE_LOOP) {
            // the workgroups' time on units, as a share of the whole GPU
            uint64_t u[3];
            double b[3];
            (void)voss_serve_unit_stats(s, u, b);
            *device_busy_s = (b[0] + b[1] + b[2]) / std::max(1, s->c->num_cus);
        } else {
            std::lock_guard<std::mutex> g(s->stat_mu);
            *device_busy_s = s->busy_ms / 1e3;
        }
    }
    if (wall_s) *wall_s = now - s->t_start;
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
    if (s->tune_saved) s->c->tune = s->saved_tune;
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
