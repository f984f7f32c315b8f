// serve_host.h -- host side of the server (voss_serve_*): the request
// record, the handle and the buffers kept by the context, shared by the
// device-resident loop's scheduler (serve.hip) and the host-stepped
// scheduler (serve_steps.hip).  Not a public header.
#pragma once
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <thread>
#include <unordered_map>

#include "engine.h"
#include "range_alloc.h"
#include "serve_launch.h"

namespace voss {
namespace serving {

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
    size_t next = 0;     // host steps: windows handed to steps so far
    int slot = -1;
    uint32_t epoch = 0;  // host steps: bumped by a restart; steps planned before it are ignored
    bool needs_init = true;  // host steps: the slot's accumulators must be cleared first
    JobHost J;
    uint64_t carry_last = 0;
    bool carry_has = false;
    voss_interval_result res{};
    std::vector<voss_large_gap> large;
    int status = VOSS_OK;
    bool done = false;
    // device loop: generation published in the slot's mailbox, arena
    // ranges (offset, length) and the sizes that a restart enlarges
    uint32_t gen = 0;
    uint64_t ent_off = 0, ent_n = 0;    // bucket entries
    uint64_t rec_off = 0, rec_n = 0;    // window records
    uint64_t cur_off = 0, cur_n = 0;    // bucket cursors
    uint64_t tile_off = 0, tile_n = 0;  // tile counters
    uint64_t host_off = 0, host_n = 0;  // pinned bytes
    uint32_t room_pct = 0;  // 0: the context's tuning
    uint64_t lcap = 0;      // 0: estimated at admission
    uint64_t dev_bytes = 0; // device bytes counted against the HBM budget
    std::vector<ServeUnit> units;  // MR chunks, or an interval's bucket tasks (tile order)
    size_t upub = 0;               // how many of them are published
    uint64_t nwin = 0, wpub = 0;   // an interval's windows, and how many are published
    std::vector<uint32_t> win_tile;      // tile of each window, ~0 without buckets
    std::vector<uint32_t> tile_end;      // units (bucket tasks) up to and including each tile's
    bool early = false;                  // a small request admitted onto an idle device: a tile's
                                         // windows are published once its tasks are (they wait
                                         // for them on the device)
    const uint32_t* tile_ready = nullptr;  // host flags the device sets per tile
    uint32_t bypassed = 0;         // later requests admitted while this one waited
};

using ReqPtr = std::shared_ptr<ServeReq>;

// One in-flight sieve step of the host-stepped scheduler: its buffers, plan,
// and which request each slot held (and at which epoch) when it was planned.
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

// One in-flight Miller-Rabin launch of the host-stepped scheduler.
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

// Pinned host memory that the device reads and writes while the loop runs
// (fine-grained: hipHostMallocCoherent).
struct PinnedBuf {
    void* p = nullptr;
    size_t bytes = 0;
    hipError_t reserve(size_t n);
    void release();
};

// Buffers of the device-resident loop.  Allocated while no loop kernel
// runs; every device counter and cursor is zero while free.
struct LoopBufs {
    uint32_t max_active = 0, mr_slots = 0;
    uint64_t ring_cap = 0;
    PinnedBuf ring, ctl, req, mail, mrq, mr_mail, host, claimed, wgstat;
    double clock_hz = 1e8;  // the device wall clock of the unit timings
    DevBuf<unsigned long long> d_ctr;  // [0] claim, [1] units done
    DevBuf<ServeSlot> d_slot;
    DevBuf<uint32_t> d_mr_done;
    DevBuf<JobAcc> d_acc;
    DevBuf<uint32_t> d_ent;
    DevBuf<SegRec> d_rec;
    DevBuf<uint32_t> d_cur;
    DevBuf<uint32_t> d_tile;
    hipEvent_t k_start = nullptr, k_done = nullptr;
    bool ready = false;
    void release();
};

// The buffers, events and streams of the schedulers.  They belong to the
// context and outlive a server, so a later server on the same context
// starts with buffers that earlier requests have grown (growing a device
// or pinned buffer synchronises the device).
struct ServeCache {
    Step steps[2];
    MrLaunch mr[2];
    DevBuf<JobAcc> d_acc;   // host steps: per-slot accumulators
    HostBuf<JobAcc> h_acc;  // host steps: read-back of finished slots
    hipStream_t mr_stream = nullptr;
    LoopBufs loop;
};

void free_serve_cache(void* p);

}  // namespace serving
}  // namespace voss

struct voss_server {
    voss_ctx* c = nullptr;
    voss_serve_config cfg{};
    std::thread th;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<voss::serving::ReqPtr> queue;
    std::unordered_map<uint64_t, voss::serving::ReqPtr> all;
    uint64_t next_id = 1;
    bool stopping = false;
    int err = VOSS_OK;
    std::string errmsg;
    // owned by the scheduler thread
    std::vector<voss::serving::ReqPtr> active;
    std::vector<int> free_slots;
    uint64_t hbm_used = 0;
    voss::serving::ServeCache* k = nullptr;  // the context's scheduler buffers
    uint64_t step_seq = 0;
    uint64_t lcap = 0;      // host steps: large-gap list capacity of a step (grows on overflow)
    bool rooms_exact = false;  // host steps: set by a bucket overflow: exact rooms from then on
    // statistics (voss_serve_stats)
    std::atomic<uint64_t> arrivals{0};  // submissions so far (wakes the collector)
    std::atomic<uint64_t> n_steps{0};
    std::atomic<uint64_t> n_windows{0};
    double busy_ms = 0;      // device time of the sieve steps / finished loop launches
    double t_start = 0;
    std::mutex stat_mu;
    // the device loop sieves 131072-byte windows: the context's tuning while
    // the server runs, restored at stop
    voss_tuning saved_tune{};
    bool tune_saved = false;
};

namespace voss {
namespace serving {

uint64_t req_hbm(const voss_server* s, const ServeReq& r);
void serve_fail_all(voss_server* s, int code, const std::string& msg);
// Marks the requests done (latency stamp, HBM and slot back) and wakes waiters.
void complete(voss_server* s, const std::vector<ReqPtr>& finished);
// The interval's result from its accumulated sums (r->J).
void finish_request(voss_server* s, ServeReq* r);

// The schedulers' thread functions.
void steps_loop(voss_server* s);
void device_loop(voss_server* s);

}  // namespace serving
}  // namespace voss
