// engine.h -- internal interfaces of the host engine, shared by api.hip
// (context, planner, executor and the C ABI entry points) and serve.hip
// (the serving scheduler).  Not a public header: the C ABI is voss/voss.h.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <condition_variable>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "call_gate.h"
#include "mr_launch.h"
#include "sieve_launch.h"
#include "voss/voss.h"
#include "wheel.h"

struct voss_ctx;

namespace voss {
namespace core {

extern thread_local std::string g_err;

int fail(int code, const std::string& msg);

int hip_fail(hipError_t e, const char* what);

#define HIPCHK(call)                                         \
    do {                                                     \
        hipError_t _e = (call);                              \
        if (_e != hipSuccess) return hip_fail(_e, #call);    \
    } while (0)

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t cap = 0;  // elements
    // grow-only; a buffer that grows again takes half as much more than
    // asked, so a run of slowly growing requests does not reallocate (a
    // device-wide synchronisation) each time
    hipError_t reserve(size_t n) {
        if (n <= cap) return hipSuccess;
        const size_t want = std::max<size_t>({n, (size_t)64, cap ? cap + cap / 2 : 0});
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        hipError_t e = hipMalloc((void**)&p, want * sizeof(T));
        if (e != hipSuccess && want > n) {
            (void)hipGetLastError();
            e = hipMalloc((void**)&p, n * sizeof(T));
            if (e == hipSuccess) cap = n;
            return e;
        }
        if (e == hipSuccess) cap = want;
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

// Grow-only pinned host staging buffer (hipHostMalloc).
template <typename T>
struct HostBuf {
    T* p = nullptr;
    size_t cap = 0;
    hipError_t reserve(size_t n) {
        if (n <= cap) return hipSuccess;
        const size_t want = std::max<size_t>({n, (size_t)64, cap ? cap + cap / 2 : 0});
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
        hipError_t e = hipHostMalloc((void**)&p, want * sizeof(T), hipHostMallocDefault);
        if (e != hipSuccess && want > n) {
            (void)hipGetLastError();
            e = hipHostMalloc((void**)&p, n * sizeof(T), hipHostMallocDefault);
            if (e == hipSuccess) cap = n;
            return e;
        }
        if (e == hipSuccess) cap = want;
        return e;
    }
    void release() {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
    }
};

// Buffers of one execution of a plan, device and pinned host.  The context
// owns one set for its calls; a prepared batch owns its own, so its
// captured graph keeps valid pointers between replays.
struct ExecBufs {
    DevBuf<SegDesc> d_desc;
    DevBuf<SegRec> d_rec;
    DevBuf<JobAcc> d_acc;
    DevBuf<LargeRaw> d_large;
    DevBuf<unsigned long long> d_flags;  // [0] large-gap records, [1] bucket overflow,
                                         // [2] bucket task counter, [3] sieve
                                         // window counter (low 32 bits)
    DevBuf<BucketTask> d_btask;
    DevBuf<uint32_t> d_bcnt;
    DevBuf<uint32_t> d_bcur;
    DevBuf<unsigned long long> d_boff;
    DevBuf<uint32_t> d_bent;
    HostBuf<SegDesc> h_desc;
    HostBuf<BucketTask> h_btask;
    HostBuf<unsigned long long> h_boff;
    HostBuf<uint32_t> h_bcnt;
    HostBuf<SegRec> h_rec;
    HostBuf<JobAcc> h_acc;
    HostBuf<LargeRaw> h_large;
    HostBuf<unsigned long long> h_flags;
    void release() {
        d_desc.release(); d_rec.release(); d_acc.release(); d_large.release();
        d_flags.release(); d_btask.release(); d_bcnt.release(); d_bcur.release();
        d_boff.release(); d_bent.release(); h_desc.release(); h_btask.release();
        h_boff.release(); h_bcnt.release(); h_rec.release(); h_acc.release();
        h_large.release(); h_flags.release();
    }
};

constexpr uint32_t kW30Mask[8] = {1, 7, 11, 13, 17, 19, 23, 29};

uint64_t isqrt_u64(uint64_t n);

// Primes 7 <= p <= limit (segmented odd-only sieve; limit < 2^32).
std::vector<uint32_t> make_base_primes(uint64_t limit);

float merit_of(uint64_t p_before, uint32_t g);

bool merit_ge10(uint64_t p_before, uint32_t g);

// A fixed set of host worker threads, started on first use, for the host
// steps of a run (planning, accumulation, merits): starting threads per
// call cost more than the sub-millisecond steps of a small run.  One
// caller at a time; a second concurrent caller, or a call from inside a
// worker, runs its work inline.
class HostPool {
public:
    static HostPool& get() {
        static HostPool pool;
        return pool;
    }
    size_t workers() const { return th_.size(); }
    // job(t) for t in [0, n), the calling thread taking part; false if the
    // pool is busy (the caller then does the work itself)
    bool run(size_t n, const std::function<void(size_t)>& job) {
        if (in_worker() || !busy_.try_lock()) return false;
        {
            std::lock_guard<std::mutex> g(m_);
            job_ = &job;
            n_ = n;
            next_ = 0;
            done_ = 0;
            gen_++;
        }
        cv_.notify_all();
        in_worker() = true;  // a nested call from one of our jobs runs inline
        work();
        in_worker() = false;
        {
            std::unique_lock<std::mutex> g(m_);
            done_cv_.wait(g, [&] { return done_ == n_; });
            job_ = nullptr;
        }
        busy_.unlock();
        return true;
    }

private:
    HostPool() {
        const size_t hw = std::max<unsigned>(1u, std::thread::hardware_concurrency());
        const size_t nw = std::min<size_t>(15, hw > 1 ? hw - 1 : 0);
        for (size_t i = 0; i < nw; i++) th_.emplace_back([this] { loop(); });
    }
    ~HostPool() {
        {
            std::lock_guard<std::mutex> g(m_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto& t : th_) t.join();
    }
    static bool& in_worker() {
        static thread_local bool flag = false;
        return flag;
    }
    // takes job indices until none is left
    void work() {
        for (;;) {
            size_t t;
            const std::function<void(size_t)>* job;
            {
                std::lock_guard<std::mutex> g(m_);
                if (!job_ || next_ >= n_) return;
                t = next_++;
                job = job_;
            }
            (*job)(t);
            std::lock_guard<std::mutex> g(m_);
            if (++done_ == n_) done_cv_.notify_all();
        }
    }
    void loop() {
        in_worker() = true;
        uint64_t seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> g(m_);
                cv_.wait(g, [&] { return stop_ || gen_ != seen; });
                if (stop_) return;
                seen = gen_;
            }
            work();
        }
    }
    std::vector<std::thread> th_;
    std::mutex m_, busy_;
    std::condition_variable cv_, done_cv_;
    const std::function<void(size_t)>* job_ = nullptr;
    size_t n_ = 0, next_ = 0, done_ = 0;
    uint64_t gen_ = 0;
    bool stop_ = false;
};

// f(begin, end) over [0, n) split into up to 16 chunks of at least
// kMinPerThread items, run on the host pool.
template <typename F>
void parallel_for(size_t n, F&& f, size_t kMinPerThread = 8192) {
    size_t nt = std::min<size_t>(16, HostPool::get().workers() + 1);
    nt = std::min(nt, std::max<size_t>(1, n / kMinPerThread));
    if (nt <= 1) {
        f((size_t)0, n);
        return;
    }
    const std::function<void(size_t)> job = [&](size_t t) { f(n * t / nt, n * (t + 1) / nt); };
    if (!HostPool::get().run(nt, job))
        for (size_t t = 0; t < nt; t++) job(t);
}

// An allocator whose resize() leaves new elements default-initialised
// (uninitialised for plain structs), so that a window list of 1.2 GB is not
// zero-filled before the planner writes every element.
template <typename T>
struct NoInitAlloc : std::allocator<T> {
    template <typename U>
    struct rebind {
        using other = NoInitAlloc<U>;
    };
    NoInitAlloc() = default;
    template <typename U>
    NoInitAlloc(const NoInitAlloc<U>&) {}
    template <typename U>
    void construct(U* p) {
        ::new (static_cast<void*>(p)) U;
    }
    template <typename U, typename... Args>
    void construct(U* p, Args&&... args) {
        ::new (static_cast<void*>(p)) U(std::forward<Args>(args)...);
    }
};

using DescVec = std::vector<SegDesc, NoInitAlloc<SegDesc>>;

using LargeVec = std::vector<voss_large_gap, NoInitAlloc<voss_large_gap>>;

// memcpy on several host threads for large blocks (pinned staging of the
// window descriptors: 1.2 GB for a plain run to 1e14).
void par_copy(void* dst, const void* src, size_t bytes);

}  // namespace core
}  // namespace voss

using namespace voss;
using namespace voss::core;

struct voss_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr;
    std::string name;
    int num_cus = 0;
    double total_gb = 0;
    voss_tuning tune{131072, 1, 768, 1u << 20, 1, 100, {0, 0}};

    std::vector<uint32_t> primes;
    std::vector<double> inv_prefix;  // inv_prefix[i] = sum of 1/p over primes[0..i)
    uint64_t primes_limit = 0;
    DevBuf<PrimeRec> d_primes;
    DevBuf<uint32_t> d_p32;  // the same primes as 32-bit values (bucket walk mode)
    size_t d_nprimes = 0;

    DevBuf<uint32_t> d_pat;
    uint32_t pat_off[kPresieveGroups] = {};
    uint32_t pat_len[kPresieveGroups] = {};
    uint32_t restore0 = 0;

    std::vector<unsigned long long> lg_thresh;
    DevBuf<unsigned long long> d_lg_thresh;

    ExecBufs eb;                  // buffers of the context's own runs
    // window list of the last plain run, kept so the next plan reuses its
    // memory (a fresh 1 MB vector per call costs page faults)
    DescVec desc_pool;
    std::vector<hipEvent_t> evs;  // per-round timing events
    int graph_mode = 1;           // batched runs as one captured hipGraph
    std::atomic<bool> served{false};  // a running server's scheduler owns it

    // per-window prime counts of the last plain run (gap-sequence sampling)
    uint64_t last_N = 0, last_seg = 0;
    uint32_t last_W = 0;
    std::vector<uint64_t> last_counts;
    DevBuf<unsigned long long> d_emit;
    DevBuf<unsigned long long> d_emit_off;

    DevBuf<uint64_t> d_mr_n;
    DevBuf<uint8_t> d_mr_v;
    DevBuf<uint32_t> d_surv;
    DevBuf<uint32_t> d_nsurv;
    uint32_t* mr_scratch = nullptr;  // the no-allocation MR entry point's lists
                                     // and counters (device address, fixed)
    DevBuf<uint64_t> d_vf_pa;   // verify_gaps: p_after, gap, ok
    DevBuf<int32_t> d_vf_gap;
    DevBuf<uint8_t> d_vf_ok;
    // buffers of the serving scheduler (serve.hip), kept between servers on
    // this context so that a new server starts with grown buffers
    void* serve_cache = nullptr;
    void (*free_serve_cache)(void*) = nullptr;
};


namespace voss {
namespace core {

// The process-wide gate of library calls (call_gate.h), and whether the
// calling thread is a server's scheduler.
CallGate& gate();

extern thread_local bool t_scheduler;

// Admission of one entry point: VOSS_EBUSY when another thread is inside the
// library, or when the context belongs to a running server (to anyone but
// that server's scheduler thread).  wait = true blocks for the gate instead
// (context setup and teardown).
struct Admit {
    int rc = VOSS_OK;
    bool held = false;
    explicit Admit(const voss_ctx* c, bool wait = false) {
        if (c && c->served.load() && !t_scheduler) {
            rc = fail(VOSS_EBUSY, "the context belongs to a running server (submit requests to it)");
            return;
        }
        if (wait) {
            gate().enter();
        } else if (!gate().try_enter()) {
            rc = fail(VOSS_EBUSY,
                      "another thread is inside the library: one call at a time per process");
            return;
        }
        held = true;
    }
    ~Admit() {
        if (held) gate().leave();
    }
    Admit(const Admit&) = delete;
    Admit& operator=(const Admit&) = delete;
};

#define VOSS_ADMIT(ctx)          \
    Admit admit_(ctx);           \
    if (admit_.rc) return admit_.rc

int set_device(voss_ctx* c);

// Base primes covering every window that ends at or below max_number.
int ensure_primes(voss_ctx* c, uint64_t max_number);

uint32_t primes_upto(const voss_ctx* c, uint64_t max_number);

// Presieve tables: for each group of three primes, one word per residue
// of the word index modulo the product (a period of the group's pattern).
int build_patterns(voss_ctx* c);

// Largest p_before for which float merit would reach 10 at gap g:
// (double)g / log((double)p) >= 10.0 is monotone in p, so binary search.
int build_lg_thresh(voss_ctx* c);

// Host-side accumulation of one job (the device accumulators plus the pairs
// joined on the host: the head of the range and launch boundaries), in the
// compact bin layout: hist2[h] = gap 2h, gap1 = gap 1 (only (2,3)).
struct JobHost {
    int64_t gap1 = 0;
    std::vector<int64_t> hist2;  // [kHistBins]
    std::vector<int64_t> mq2;    // [kModqRows * kModqBins]
    uint64_t count = 0, c1 = 0, mod30[8] = {};
    uint64_t first = 0, last = 0;
    uint64_t large_count = 0;
    bool overflow = false;       // a gap >= 2048 (none exists below 2^64)
    LargeVec large;
    // the histograms are allocated on first use (init), so that a batch's
    // thousands of jobs can be set up on several host threads
    void init() {
        if (hist2.empty()) {
            hist2.assign(kHistBins, 0);
            mq2.assign(kModqRows * kModqBins, 0);
        }
    }
};

void host_pair(JobHost& J, uint64_t p, uint64_t q, bool full);

void add_acc(JobHost& J, const JobAcc& A);

// Canonical order of the large-gap list: merit descending, then position
// ascending (the reference sorts by merit alone with an unstable sort,
// src/voss_v7_live.cu:813, so ties have no defined order there).  Merits
// are positive floats, whose bit patterns order like the values: an LSD
// radix sort (11-bit digits, digits that never vary skipped) of the
// inverted merit bits carrying the list index, then runs of equal merit
// sorted by position.
void sort_large(LargeVec& v);

// The canonical order of sort_large as a comparison.
bool large_before(const voss_large_gap& x, const voss_large_gap& y);

// The k first records of the canonical order into dst (k <= v.size()):
// each host thread selects its chunk's best k, then the best k of those.
// For long lists of which the caller keeps a few (a plain run to 1e14 has
// 4.2e7 candidates).
void top_large(const LargeVec& v, size_t k, voss_large_gap* dst);

// Histograms and large gaps into the reference's array layout.
void write_result(JobHost& J, voss_result* out);

// Wheel-prime contribution to the residue counters (slot layout of the
// reference's modq_d, src/voss_v7_live.cu:195-232).
void wheel_modq(const JobHost& J, uint64_t* modq);

struct Plan {
    DescVec desc;
    uint32_t njobs = 0;
};

// Windows of the grid w0 = k*W covering numbers [a, b] (a >= 7): k from
// ka = (a/30 - origin)/W to kb = (b/30 - origin)/W, each window non-empty
// (30*w0 <= b and 30*(w0+W) > a for every such k).
size_t plan_count(uint64_t a, uint64_t b, uint32_t W, uint64_t grid_origin);

// Writes the plan_count(a, b, W, origin) windows of [a, b] to out.  The
// base-prime count of consecutive windows only grows, so it is advanced
// with one pointer instead of a search per window.
void plan_fill(const voss_ctx* c, SegDesc* out, uint64_t a, uint64_t b, uint32_t job,
               uint32_t lseg, uint32_t W, uint64_t grid_origin);

// Appends the windows of [a, b] to P.
void plan_range(const voss_ctx* c, Plan& P, uint64_t a, uint64_t b, uint32_t job,
                uint32_t lseg, uint32_t W, uint64_t grid_origin);

SieveArgs make_args(voss_ctx* c, ExecBufs& B, uint32_t ndesc, unsigned long long large_cap);

constexpr size_t kMaxDescPerLaunch = 1u << 20;

struct ExecTimes {
    double device_ms = 0;  // the whole device sequence
    double sieve_ms = 0;   // sieve kernels (-1: not timed, under a graph)
    double bucket_ms = 0;  // bucket kernels (-1: not timed, under a graph)
    uint64_t bucket_entries = 0;
    uint32_t rounds = 0;
    uint32_t reruns = 0;   // re-executions after a bucket or large-list overflow
};

// Bucket entries per round (32 GiB of HBM).  Bucket offsets are 64-bit; a
// fill task's slots are 32-bit offsets from its first bucket.  Larger
// rounds hold more tiles at once, so fewer of them are split into prime
// chunks that count before they write (4096-interval batch, bucket fills:
// 90.7 ms at 2^31 entries per round, 83.4 ms at 2^33).
constexpr uint64_t kBucketEntryBudget = 1ull << 33;

// Room for the entries of one bucket (a window of W bytes) fed by primes
// with sum(1/p) = s.  A prime p has on average 8W/p multiples coprime to 30
// among the window's 30W numbers; the window's count is a sum of nearly
// independent per-prime counts whose variance is at most their mean, so
// mean + 8 sigma + 64 is essentially never exceeded.  If it is, the fill
// kernel reports it and the plan is re-run with exact counts.
uint32_t bucket_room(double s, uint32_t W, uint32_t pct);

// A round is a run of consecutive windows of the plan launched together:
// bucket fill, sieve, stitch.  Windows of jobs with at least
// bucket_min_windows windows get a bucket for the primes >= bucket_limit.
struct Round {
    size_t off = 0, nd = 0;        // plan windows [off, off + nd)
    DescVec desc;                  // with bucket / nprimes set for this round
    std::vector<BucketTask> tasks;
    uint32_t nbuckets = 0;
    std::vector<uint32_t> room;    // entries per bucket
    size_t task_off = 0, bkt_off = 0, boff_off = 0;  // places in the staged arrays
    uint64_t room_total = 0;
};

// Estimated VALU work of a bucket task: per-prime setup (a class-mode
// prime also starts eight classes), class-mode hits, and walk steps (30
// span/p per walked prime).
double bucket_task_cost(const voss_ctx* c, const BucketTask& T, uint32_t W);

// min_share: a tile is split into shared tasks only beyond this estimated
// cost (and beyond a fair share of the whole plan).
void make_rounds(const voss_ctx* c, const Plan& P, std::vector<Round>& rounds,
                 double min_share = 0.0);

int log2_window(const voss_ctx* c);

// Runs a plan on the device: every round's bucket fill, sieve and stitch
// in stream order, then the accumulators, window records and flags back
// to pinned memory, with a single synchronisation at the end.  The direct
// sequence has no host step inside, so it can be captured in a hipGraph
// and replayed.  Each job's accumulators sum over all its rounds on the
// device; the pairs across round edges are joined on the host from the
// window records.
struct Executor {
    voss_ctx* c = nullptr;
    ExecBufs* B = nullptr;
    Plan P;
    std::vector<Round> rounds;
    size_t ntask = 0, nbkt = 0, nboff = 0;
    uint64_t lcap = 0;         // device large-gap list capacity
    bool staged = false;       // buffers sized and staging arrays filled
    bool exact = false;        // bucket rooms are the exact counts
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    const PrimeRec* graph_primes = nullptr;  // base-prime tables the graph reads
    const uint32_t* graph_p32 = nullptr;
    bool copy_acc = true;      // read the accumulators back after a run
    uint64_t nlarge = 0;       // large-gap records of the last run (in B->h_large)
    // Miller-Rabin batches run beside the rounds on a second stream (forked
    // from and joined back into the main one, so a capture holds both)
    struct MrJob {
        const uint64_t* d_n;
        uint64_t count;
        uint8_t* d_verdict;
        uint32_t* d_surv;
        uint32_t* d_nsurv;
    };
    std::vector<MrJob> mr;
    hipStream_t side = nullptr;
    hipEvent_t fork_ev = nullptr, join_ev = nullptr;

    Executor(voss_ctx* ctx, ExecBufs* bufs) : c(ctx), B(bufs) {}
    ~Executor() {
        drop_graph();
        if (fork_ev) (void)hipEventDestroy(fork_ev);
        if (join_ev) (void)hipEventDestroy(join_ev);
        if (side) (void)hipStreamDestroy(side);
    }

    int ensure_side() {
        if (side) return VOSS_OK;
        HIPCHK(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
        HIPCHK(hipEventCreateWithFlags(&fork_ev, hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&join_ev, hipEventDisableTiming));
        return VOSS_OK;
    }

    // fork: the side stream waits for the main one, runs the MR batches
    int enqueue_mr_fork() {
        if (mr.empty()) return VOSS_OK;
        HIPCHK(hipEventRecord(fork_ev, c->stream));
        HIPCHK(hipStreamWaitEvent(side, fork_ev, 0));
        for (const MrJob& m : mr)
            HIPCHK(launch_mr(m.d_n, m.count, m.d_verdict, m.d_surv, m.d_nsurv, c->num_cus, side));
        return VOSS_OK;
    }

    // join: the main stream waits for the side stream
    int enqueue_mr_join() {
        if (mr.empty()) return VOSS_OK;
        HIPCHK(hipEventRecord(join_ev, side));
        HIPCHK(hipStreamWaitEvent(c->stream, join_ev, 0));
        return VOSS_OK;
    }
    Executor(const Executor&) = delete;
    Executor& operator=(const Executor&) = delete;

    void drop_graph() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        gexec = nullptr;
        graph = nullptr;
    }

    void plan_rounds() {
        rounds.clear();
        make_rounds(c, P, rounds);
        ntask = nbkt = nboff = 0;
        for (Round& R : rounds) {
            R.task_off = ntask;
            R.bkt_off = nbkt;
            R.boff_off = nboff;
            ntask += R.tasks.size();
            nbkt += R.nbuckets;
            nboff += R.nbuckets + 1;
        }
        staged = false;
        exact = false;
        drop_graph();
    }

    // Sizes every buffer and fills the pinned staging arrays (descriptors,
    // tasks, bucket offsets from the rooms).  Nothing is allocated after
    // this until the next staging, so captured pointers stay valid.
    int stage() {
        const size_t nd = P.desc.size();
        const uint32_t nj = std::max<uint32_t>(P.njobs, 1);
        if (lcap == 0) {
            // room for every large-gap candidate: about 1.5e-5 of the primes
            // at 1e11..1e14 (8.9e-6 at 1e11, 1.3e-5 at 1e13); 2e-5 of a
            // prime-count estimate avoids a re-run for overflow in practice
            // (a sample of at most ~65536 windows, scaled)
            const size_t step = std::max<size_t>(16, P.desc.size() >> 16);
            double primes = 0;
            for (size_t i = 0; i < P.desc.size(); i += step) {
                const SegDesc& d = P.desc[i];
                primes += (double)step * (double)(d.n_hi - d.n_lo + 1) /
                          std::log((double)d.n_hi + 3.0);
            }
            lcap = std::max<uint64_t>({B->d_large.cap, 1u << 20, (uint64_t)(2e-5 * primes)});
        }
        uint64_t ent = 1;
        uint32_t nb_max = 1;
        for (const Round& R : rounds) {
            ent = std::max<uint64_t>(ent, R.room_total);
            nb_max = std::max<uint32_t>(nb_max, R.nbuckets);
        }
        if (ent > 2 * kBucketEntryBudget)
            return fail(VOSS_ENOMEM, "bucket round beyond its entry budget");
        drop_graph();
        HIPCHK(B->d_desc.reserve(nd));
        HIPCHK(B->d_rec.reserve(nd));
        HIPCHK(B->d_acc.reserve(nj));
        HIPCHK(B->d_large.reserve(lcap));
        HIPCHK(B->d_flags.reserve(4));
        HIPCHK(B->d_btask.reserve(ntask));
        HIPCHK(B->d_bcnt.reserve(nb_max));
        HIPCHK(B->d_bcur.reserve(nbkt));
        HIPCHK(B->d_boff.reserve(nboff));
        HIPCHK(B->d_bent.reserve(ent));
        HIPCHK(B->h_desc.reserve(nd));
        HIPCHK(B->h_btask.reserve(ntask));
        HIPCHK(B->h_boff.reserve(nboff));
        HIPCHK(B->h_bcnt.reserve(nb_max));
        HIPCHK(B->h_rec.reserve(nd));
        HIPCHK(B->h_acc.reserve(nj));
        HIPCHK(B->h_flags.reserve(2));
        for (const Round& R : rounds) {
            par_copy(B->h_desc.p + R.off, R.desc.empty() ? P.desc.data() + R.off : R.desc.data(),
                     R.nd * sizeof(SegDesc));
            if (!R.tasks.empty())
                std::memcpy(B->h_btask.p + R.task_off, R.tasks.data(),
                            R.tasks.size() * sizeof(BucketTask));
            unsigned long long o = 0;
            for (uint32_t b = 0; b < R.nbuckets; b++) {
                B->h_boff.p[R.boff_off + b] = o;
                o += R.room[b];
            }
            B->h_boff.p[R.boff_off + R.nbuckets] = o;
        }
        staged = true;
        return VOSS_OK;
    }

    hipEvent_t ev(size_t i) { return c->evs[i]; }

    uint32_t* task_counter() { return reinterpret_cast<uint32_t*>(B->d_flags.p + 2); }
    uint32_t* win_counter() { return reinterpret_cast<uint32_t*>(B->d_flags.p + 3); }

    int ensure_events(size_t n) {
        while (c->evs.size() < n) {
            hipEvent_t e;
            HIPCHK(hipEventCreate(&e));
            c->evs.push_back(e);
        }
        return VOSS_OK;
    }

    SieveArgs round_args(const Round& R) {
        SieveArgs a = make_args(c, *B, (uint32_t)R.nd, lcap);
        a.desc = B->d_desc.p + R.off;
        a.rec = B->d_rec.p + R.off;
        a.bucket_off = B->d_boff.p + R.boff_off;
        a.bucket_cnt = B->d_bcur.p + R.bkt_off;
        return a;
    }

    // The direct sequence (no host step): capturable.  timed records four
    // events per round (not under capture).
    int enqueue_direct(bool timed) {
        hipStream_t s = c->stream;
        const size_t nd = P.desc.size();
        const int log2w = log2_window(c);
        const uint32_t grid = (uint32_t)c->num_cus * std::max<uint32_t>(c->tune.wg_per_cu, 1);
        const uint32_t bgrid = (uint32_t)c->num_cus * 2;
        if (nd)
            HIPCHK(hipMemcpyAsync(B->d_desc.p, B->h_desc.p, nd * sizeof(SegDesc),
                                  hipMemcpyHostToDevice, s));
        if (ntask) {
            HIPCHK(hipMemcpyAsync(B->d_btask.p, B->h_btask.p, ntask * sizeof(BucketTask),
                                  hipMemcpyHostToDevice, s));
            HIPCHK(hipMemcpyAsync(B->d_boff.p, B->h_boff.p, nboff * 8, hipMemcpyHostToDevice, s));
            HIPCHK(launch_fill_u32(B->d_bcur.p, nbkt, 0, s));
        }
        HIPCHK(launch_fill_u32((uint32_t*)B->d_flags.p, 4, 0, s));
        HIPCHK(launch_init_acc(B->d_acc.p, P.njobs, s));
        int rc = enqueue_mr_fork();
        if (rc) return rc;
        for (size_t r = 0; r < rounds.size(); r++) {
            const Round& R = rounds[r];
            if (timed) HIPCHK(hipEventRecord(ev(4 * r), s));
            if (!R.tasks.empty()) {
                HIPCHK(launch_fill_u32(task_counter(), 1, 0, s));
                HIPCHK(launch_bucket_fill(B->d_btask.p + R.task_off, (uint32_t)R.tasks.size(),
                                          c->d_primes.p, c->d_p32.p, B->d_boff.p + R.boff_off,
                                          B->d_bcur.p + R.bkt_off, B->d_bent.p,
                                          (uint32_t*)(B->d_flags.p + 1), task_counter(), log2w,
                                          bgrid, s));
            }
            if (timed) HIPCHK(hipEventRecord(ev(4 * r + 1), s));
            const SieveArgs a = round_args(R);
            HIPCHK(launch_fill_u32(win_counter(), 1, 0, s));
            HIPCHK(launch_sieve(a, (int)c->tune.window_bytes, grid, s));
            if (timed) HIPCHK(hipEventRecord(ev(4 * r + 2), s));
            HIPCHK(launch_stitch(a, s));
            if (timed) HIPCHK(hipEventRecord(ev(4 * r + 3), s));
        }
        if ((rc = enqueue_mr_join())) return rc;
        return enqueue_readback();
    }

    int enqueue_readback() {
        hipStream_t s = c->stream;
        if (copy_acc && P.njobs)
            HIPCHK(hipMemcpyAsync(B->h_acc.p, B->d_acc.p, P.njobs * sizeof(JobAcc),
                                  hipMemcpyDeviceToHost, s));
        if (!P.desc.empty())
            HIPCHK(hipMemcpyAsync(B->h_rec.p, B->d_rec.p, P.desc.size() * sizeof(SegRec),
                                  hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(B->h_flags.p, B->d_flags.p, 16, hipMemcpyDeviceToHost, s));
        return VOSS_OK;
    }

    // Exact sequence: per round, a count pass, the host prefix sum, then the
    // fill with exact offsets.  The rooms become the exact counts, so later
    // direct runs of this plan cannot overflow.
    int enqueue_exact(bool timed) {
        hipStream_t s = c->stream;
        const size_t nd = P.desc.size();
        const int log2w = log2_window(c);
        const uint32_t grid = (uint32_t)c->num_cus * std::max<uint32_t>(c->tune.wg_per_cu, 1);
        const uint32_t bgrid = (uint32_t)c->num_cus * 2;
        HIPCHK(hipMemcpyAsync(B->d_desc.p, B->h_desc.p, nd * sizeof(SegDesc),
                              hipMemcpyHostToDevice, s));
        if (ntask)
            HIPCHK(hipMemcpyAsync(B->d_btask.p, B->h_btask.p, ntask * sizeof(BucketTask),
                                  hipMemcpyHostToDevice, s));
        HIPCHK(launch_fill_u32((uint32_t*)B->d_flags.p, 4, 0, s));
        HIPCHK(launch_init_acc(B->d_acc.p, P.njobs, s));
        for (size_t r = 0; r < rounds.size(); r++) {
            Round& R = rounds[r];
            if (timed) HIPCHK(hipEventRecord(ev(4 * r), s));
            if (!R.tasks.empty()) {
                HIPCHK(launch_fill_u32(B->d_bcnt.p, R.nbuckets, 0, s));
                HIPCHK(launch_fill_u32(task_counter(), 1, 0, s));
                HIPCHK(launch_bucket_count(B->d_btask.p + R.task_off, (uint32_t)R.tasks.size(),
                                           c->d_primes.p, c->d_p32.p, B->d_bcnt.p, task_counter(), log2w,
                                           bgrid, s));
                HIPCHK(hipMemcpyAsync(B->h_bcnt.p, B->d_bcnt.p, R.nbuckets * 4,
                                      hipMemcpyDeviceToHost, s));
                HIPCHK(hipStreamSynchronize(s));
                unsigned long long o = 0;
                for (uint32_t b = 0; b < R.nbuckets; b++) {
                    B->h_boff.p[R.boff_off + b] = o;
                    R.room[b] = (B->h_bcnt.p[b] + 3u) & ~3u;  // 16-byte aligned buckets
                    o += R.room[b];
                }
                B->h_boff.p[R.boff_off + R.nbuckets] = o;
                R.room_total = o;
                if (o > 2 * kBucketEntryBudget)
                    return fail(VOSS_ENOMEM, "bucket round beyond its entry budget");
                // the stream is idle here: growing the entry buffer is safe
                HIPCHK(B->d_bent.reserve(std::max<uint64_t>(o, 1)));
                HIPCHK(hipMemcpyAsync(B->d_boff.p + R.boff_off, B->h_boff.p + R.boff_off,
                                      (R.nbuckets + 1) * 8, hipMemcpyHostToDevice, s));
                HIPCHK(launch_fill_u32(B->d_bcur.p + R.bkt_off, R.nbuckets, 0, s));
                HIPCHK(launch_fill_u32(task_counter(), 1, 0, s));
                HIPCHK(launch_bucket_fill(B->d_btask.p + R.task_off, (uint32_t)R.tasks.size(),
                                          c->d_primes.p, c->d_p32.p, B->d_boff.p + R.boff_off,
                                          B->d_bcur.p + R.bkt_off, B->d_bent.p,
                                          (uint32_t*)(B->d_flags.p + 1), task_counter(), log2w,
                                          bgrid, s));
            }
            if (timed) HIPCHK(hipEventRecord(ev(4 * r + 1), s));
            const SieveArgs a = round_args(R);
            HIPCHK(launch_fill_u32(win_counter(), 1, 0, s));
            HIPCHK(launch_sieve(a, (int)c->tune.window_bytes, grid, s));
            if (timed) HIPCHK(hipEventRecord(ev(4 * r + 2), s));
            HIPCHK(launch_stitch(a, s));
            if (timed) HIPCHK(hipEventRecord(ev(4 * r + 3), s));
        }
        return enqueue_readback();
    }

    int capture() {
        drop_graph();
        HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        const int rc = enqueue_direct(false);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        HIPCHK(e);
        graph = g;
        HIPCHK(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
        graph_primes = c->d_primes.p;
        graph_p32 = c->d_p32.p;
        return VOSS_OK;
    }

    // Executes the plan until it completes without overflow.  On return
    // B->h_acc, B->h_rec and B->h_large (nlarge records) hold the results.
    int run(bool use_graph, ExecTimes* tm) {
        ExecTimes t;
        t.rounds = (uint32_t)rounds.size();
        int rc;
        if ((rc = ensure_events(4 * rounds.size()))) return rc;
        if (gexec && (graph_primes != c->d_primes.p || graph_p32 != c->d_p32.p))
            drop_graph();  // base primes moved
        bool breakdown = false;
        for (;;) {
            if (!staged && (rc = stage())) return rc;
            const bool timed = !use_graph;
            HIPCHK(hipEventRecord(c->ev0, c->stream));
            if (use_graph) {
                if (!gexec && (rc = capture())) return rc;
                HIPCHK(hipGraphLaunch(gexec, c->stream));
            } else if ((rc = enqueue_direct(timed))) {
                return rc;
            }
            HIPCHK(hipEventRecord(c->ev1, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
            breakdown = timed;
            if (B->h_flags.p[1]) {
                // a bucket outgrew its room: exact counts, then direct again
                t.reruns++;
                HIPCHK(hipEventRecord(c->ev0, c->stream));
                breakdown = true;
                if ((rc = enqueue_exact(true))) return rc;
                HIPCHK(hipEventRecord(c->ev1, c->stream));
                HIPCHK(hipStreamSynchronize(c->stream));
                staged = false;  // re-stage with the exact rooms next time
                exact = true;
                if (B->h_flags.p[1]) return fail(VOSS_EDEVICE, "bucket overflow with exact counts");
            }
            if (B->h_flags.p[0] > lcap) {
                // keep every candidate (canonical top-K needs all of them)
                t.reruns++;
                lcap = B->h_flags.p[0] + (B->h_flags.p[0] >> 2);
                staged = false;
                continue;
            }
            break;
        }
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
        t.device_ms = ms;
        if (!breakdown) {
            t.sieve_ms = t.bucket_ms = -1;
        } else {
            for (size_t r = 0; r < rounds.size(); r++) {
                float b = 0, sv = 0;
                HIPCHK(hipEventElapsedTime(&b, ev(4 * r), ev(4 * r + 1)));
                HIPCHK(hipEventElapsedTime(&sv, ev(4 * r + 1), ev(4 * r + 2)));
                t.bucket_ms += b;
                t.sieve_ms += sv;
            }
        }
        for (const Round& R : rounds) t.bucket_entries += R.room_total;
        nlarge = B->h_flags.p[0];
        HIPCHK(B->h_large.reserve(std::max<uint64_t>(nlarge, 1)));
        if (nlarge) {
            HIPCHK(hipMemcpyAsync(B->h_large.p, B->d_large.p, nlarge * sizeof(LargeRaw),
                                  hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
        }
        if (tm) *tm = t;
        return VOSS_OK;
    }

    // Folds the results into per-job host accumulators: device sums, large
    // records, and the pairs across round edges.
    int collect(std::vector<JobHost>& jobs, std::vector<uint64_t>* counts) {
        jobs.assign(P.njobs, JobHost());
        for (uint32_t j = 0; j < P.njobs; j++)
            if (B->h_acc.p[j].error) return fail(VOSS_ERANGE, "prime gap >= 2048 (beyond the device histogram)");
        parallel_for(P.njobs, [&](size_t j0, size_t j1) {
            for (size_t j = j0; j < j1; j++) add_acc(jobs[j], B->h_acc.p[j]);
        }, 256);
        // large records with their merits (one log each, on several host
        // threads for long lists), straight into the job lists: a single
        // job's list in device order, several jobs' via per-job offsets
        const LargeRaw* raw = B->h_large.p;
        auto fill = [&](voss_large_gap* dst, const uint32_t* at) {
            parallel_for(nlarge, [&](size_t k0, size_t k1) {
                for (size_t k = k0; k < k1; k++) {
                    const LargeRaw& r = raw[k];
                    dst[at ? at[k] : k] = {r.p_before + r.gap, (int32_t)r.gap,
                                           merit_of(r.p_before, r.gap)};
                }
            });
        };
        if (P.njobs == 1) {
            jobs[0].large.resize(nlarge);
            fill(jobs[0].large.data(), nullptr);
        } else if (nlarge) {
            std::vector<uint32_t> n_of(P.njobs, 0), at(nlarge);
            for (uint64_t k = 0; k < nlarge; k++) at[k] = n_of[raw[k].job]++;
            LargeVec all(nlarge);
            std::vector<uint64_t> first(P.njobs + 1, 0);
            for (uint32_t j = 0; j < P.njobs; j++) first[j + 1] = first[j] + n_of[j];
            for (uint64_t k = 0; k < nlarge; k++) at[k] += (uint32_t)first[raw[k].job];
            fill(all.data(), at.data());
            for (uint32_t j = 0; j < P.njobs; j++)
                if (n_of[j])
                    jobs[j].large.assign(all.begin() + first[j], all.begin() + first[j + 1]);
        }
        const SegRec* rec = B->h_rec.p;
        if (counts) {
            counts->resize(P.desc.size());
            parallel_for(P.desc.size(), [&](size_t i0, size_t i1) {
                for (size_t i = i0; i < i1; i++) (*counts)[i] = rec[i].count;
            }, 1u << 18);
        }
        // round edge: pair the last prime of a job before the round with the
        // first non-empty window of the same job in the round (nothing to do
        // for a single round: the stitch kernel joined all its windows)
        if (rounds.size() < 2) return VOSS_OK;
        if (P.njobs == 1) {
            // one job: only the windows next to each edge matter, found by
            // scanning from the edge (plain runs to 1e14 have 2.5e7 windows)
            bool has = false;
            uint64_t last = 0;
            uint32_t lseg = 0;
            for (const Round& R : rounds) {
                size_t i = R.off;
                const size_t e = R.off + R.nd;
                while (i < e && rec[i].count == 0) i++;
                if (i == e) continue;
                if (has) host_pair(jobs[0], last, rec[i].first, lseg == P.desc[i].lseg);
                size_t k = e - 1;
                while (rec[k].count == 0) k--;  // stops at i at the latest
                has = true;
                last = rec[k].last;
                lseg = P.desc[k].lseg;
            }
            return VOSS_OK;
        }
        std::vector<uint64_t> carry_last(P.njobs, 0);
        std::vector<uint32_t> carry_lseg(P.njobs, 0);
        std::vector<uint8_t> carry_has(P.njobs, 0), seen(P.njobs, 0);
        for (const Round& R : rounds) {
            std::fill(seen.begin(), seen.end(), 0);
            for (size_t i = R.off; i < R.off + R.nd; i++) {
                const SegDesc& d = P.desc[i];
                if (rec[i].count == 0) continue;
                if (!seen[d.job]) {
                    seen[d.job] = 1;
                    if (carry_has[d.job])
                        host_pair(jobs[d.job], carry_last[d.job], rec[i].first,
                                  carry_lseg[d.job] == d.lseg);
                }
                carry_has[d.job] = 1;
                carry_last[d.job] = rec[i].last;
                carry_lseg[d.job] = d.lseg;
            }
        }
        return VOSS_OK;
    }
};

// One-shot execution with the context's buffers.
int execute(voss_ctx* c, Plan&& P, std::vector<JobHost>& jobs, ExecTimes* times,
            std::vector<uint64_t>* counts = nullptr, bool use_graph = false);

double now_s();

int check_result(voss_result* out);

// Zeroes every field but the caller's large-gap buffer.  The two big
// histograms are left to write_result, which fills them completely.
void clear_result(voss_result* out, bool histograms = true);

// Interval bookkeeping shared by the single and batched forms: primes
// 2, 3, 5 of [lo, hi) precede the wheel primes.
struct IntervalHead {
    uint64_t total = 0, first = 0, last = 0, c1 = 0, c3 = 0;
    uint64_t modq[VOSS_Q_SLOTS] = {};
};

IntervalHead finish_interval_common(JobHost& J, uint64_t lo, uint64_t hi);

void finish_interval(JobHost& J, uint64_t lo, uint64_t hi, voss_result* out);

void finish_interval(JobHost& J, uint64_t lo, uint64_t hi, voss_interval_result* out);

}  // namespace core
}  // namespace voss
