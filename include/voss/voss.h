/*
 * voss.h -- C ABI of the MI355X (gfx950) VOSS prime-gap engine.
 *
 * The reference (src/voss_v7_live.cu, src/verify_cpp.cpp in the reference
 * tree) has no library API: its boundary is a process boundary (constants
 * patched into the CUDA source, files written into cwd, "Total primes:" on
 * stdout, exit codes).  This header is the in-process boundary underneath
 * that process boundary.  The CLI binaries voss_v7_live and verify_cpp in
 * this package are thin C++ programs over it, and src/voss_master.py drives
 * them exactly as the reference's run_voss.py expects.  Each entry point
 * names the reference code it replaces.
 *
 * Conventions
 *   - All functions return VOSS_OK (0) or a negative VOSS_E* code; none of
 *     them calls exit().  voss_last_error() gives the message of the last
 *     failure on the calling thread.
 *   - Plain pointers and sizes only.  Output buffers are caller-owned.
 *   - A voss_ctx owns device memory, streams and cached base-prime tables
 *     for one GPU.  One host thread is inside the library at a time per
 *     process: an entry point called while another thread's call runs
 *     returns VOSS_EBUSY at once, without touching the device (context
 *     creation and destruction, voss_batch_free and a server's steps wait
 *     for the running call instead).  While a server runs, its context
 *     answers direct calls with VOSS_EBUSY; clients submit requests, which
 *     is the entry point for concurrent use.
 *   - Integers only on the device; merit is computed on the host with the C
 *     library's log, as float((double)gap / log((double)p_before)).
 */
#ifndef VOSS_VOSS_H
#define VOSS_VOSS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VOSS_OK 0
#define VOSS_EINVAL (-1)  /* bad argument (range, SEG_NUM % 30, ...)        */
#define VOSS_EDEVICE (-2) /* HIP runtime / kernel failure                   */
#define VOSS_ENOMEM (-3)  /* device or host allocation failed               */
#define VOSS_ERANGE (-4)  /* a gap beyond the device histogram (never < 2^64) */
#define VOSS_EBUSY (-5)   /* another thread is inside the library, the
                             context belongs to a running server, or the
                             server is stopping                              */
#define VOSS_ENOTFOUND (-6) /* unknown serving request id                   */

/* Limits of the reference's output arrays (src/voss_v7_live.cu:23-35). */
#define VOSS_MAX_GAP 100000
#define VOSS_MODQ_MAX_GAP 100
#define VOSS_MODQ_ROWS 10 /* q=6 residues {1,5}, then q=30 residues {1,7,...,29} */
#define VOSS_Q_SLOTS 40   /* q=6 at slots 0..5, q=30 at slots 6..35 */
#define VOSS_GAP_SEQ_SIZE 5000000

/* struct LargeGap (src/voss_v7_live.cu:40): position is p_after. */
typedef struct {
    uint64_t position;
    int32_t gap;
    float merit;
} voss_large_gap;

/* Result of one sieve run.  Plain mode follows the reference's conventions
 * exactly (total_primes includes 2, 3, 5; c1/c3/modq count primes >= 7;
 * gaps[1] and gaps[2] include (2,3) and (3,5); modq_gap and large omit
 * pairs that straddle a SEG_NUM boundary).  Interval mode counts every
 * prime of [lo, hi) by its true residue and every consecutive pair inside
 * the interval. */
typedef struct {
    uint64_t total_primes;
    uint64_t first_prime; /* 0 if the range holds no prime */
    uint64_t last_prime;  /* plain mode: 5 when N < 7, like the reference */
    uint64_t c1, c3;      /* primes with p%4==1 / p%4==3 (see above) */
    uint64_t modq[VOSS_Q_SLOTS];
    int64_t gaps[VOSS_MAX_GAP];
    int64_t modq_gap[VOSS_MODQ_ROWS * VOSS_MODQ_MAX_GAP];
    uint64_t large_count;    /* out: untruncated number of large gaps */
    uint64_t large_capacity; /* in: entries available at `large` */
    voss_large_gap* large;   /* in: caller buffer or NULL; out: sorted by
                                merit desc, then position asc */
    double seconds;          /* out: wall time from planning to the final
                                histogram on the host (src/voss_v7_live.cu:
                                432 -> 632 window) */
    double device_ms;        /* out: device time of the whole run: uploads,
                                bucket fills, sieves, stitches, read-backs
                                (HIP events) */
    double sieve_ms;         /* out: device time of the sieve kernels alone
                                (-1 when the run was one captured graph) */
    uint64_t windows;        /* out: LDS windows sieved */
    double bucket_ms;        /* out: device time of the large-prime bucket
                                fills (-1 under a graph) */
    uint64_t bucket_entries; /* out: bucket room reserved for large-prime hits */
} voss_result;

/* Compact result of one interval of a batch.  gaps[h] counts the gap 2h
 * (h >= 1) and gaps[0] the gap 1 (the pair (2,3)); every prime gap below
 * 2^64 is below 2048.  modq_gap[row * VOSS_MODQ_BINS + h] counts the gap 2h
 * of the row's residue class (rows as in voss_result). */
#define VOSS_GAP_BINS 1024
#define VOSS_MODQ_BINS 50
typedef struct {
    uint64_t lo, hi;
    uint64_t total_primes;
    uint64_t first_prime;
    uint64_t last_prime;
    uint64_t c1, c3;
    uint64_t modq[VOSS_Q_SLOTS];
    uint64_t gaps[VOSS_GAP_BINS];
    uint64_t modq_gap[VOSS_MODQ_ROWS * VOSS_MODQ_BINS];
    uint64_t large_count;
    uint64_t large_capacity;
    voss_large_gap* large;
} voss_interval_result;

typedef struct voss_ctx voss_ctx;

/* Tunables chosen by voss_autotune() or by hand; zero fields keep the
 * built-in defaults. */
typedef struct {
    uint32_t window_bytes;  /* wheel bytes per LDS window: 32768, 65536, 131072 */
    uint32_t wg_per_cu;     /* resident workgroups per CU in the sieve grid */
    uint32_t tier_a_limit;  /* primes below it mark with a whole wave each */
    uint32_t bucket_limit;  /* primes >= it go through HBM bucket lists
                               (jobs of >= bucket_min_windows windows);
                               UINT32_MAX disables buckets */
    uint32_t bucket_min_windows;
    uint32_t bucket_room_pct; /* room of a bucket as a percentage of its
                                 estimate (default 100: mean + 8 sigma);
                                 a bucket that outgrows it makes the run
                                 repeat with exact counts */
    uint32_t reserved[2];
} voss_tuning;

/* ---------------------------------------------------------------- context */

/* Replaces the reference's per-run setup in main(): cudaGetDeviceProperties,
 * cudaStreamCreate and the cudaMalloc block (src/voss_v7_live.cu:353-401).
 * Calls on two contexts from two threads at once are refused (VOSS_EBUSY,
 * see the conventions above): such a run once stopped making progress on
 * an MI355X, and device-synchronising calls of one thread (hipFree,
 * hipHostFree, null-stream copies) may meet the other's stream capture.
 * The serve API is the entry point for concurrent clients. */
int voss_create(int device, voss_ctx** out);
void voss_destroy(voss_ctx* ctx);
const char* voss_last_error(void);
int voss_device_name(voss_ctx* ctx, char* buf, size_t cap, int* num_cus,
                     double* total_gb);
int voss_set_tuning(voss_ctx* ctx, const voss_tuning* t);
int voss_get_tuning(voss_ctx* ctx, voss_tuning* t);

/* Autotuner: coordinate descent over window_bytes, tier_a_limit,
 * bucket_limit and bucket_min_windows, timing voss_run_batch on the given
 * intervals (at most max_intervals of them, evenly spread) with `repeats`
 * timed runs per candidate after one warm-up; when that sample is a part of
 * the n intervals, the chosen bucket_limit and its halves/doubles are then
 * timed on all n.  Installs the fastest tuning on the context and returns
 * it in best (and in best_seconds its time on the last set it was timed
 * on).  Results never depend on the tuning, only the time does. */
int voss_autotune(voss_ctx* ctx, const uint64_t* lo, const uint64_t* hi, size_t n,
                  size_t max_intervals, int repeats, voss_tuning* best,
                  double* best_seconds);

/* ------------------------------------------------------------------ sieve */

/* Plain [1, N] run with logical SEG_NUM segments: the whole hot loop of the
 * reference's main() -- per-segment planning and sieve_w30_seg_kernel
 * (:434-502), extract_w30_seg_kernel + thrust::sort (:506-524), the host
 * boundary gap (:551-562), mod4_count_kernel / modq_count_kernel (:564-570),
 * gaps_w30_seg_kernel (:574-579), the host reduce (:581-594) and the final
 * +1 at g=1,2 (:634-635).  Requires N >= 5 and seg_num % 30 == 0, like the
 * reference (:340-342, :404). */
int voss_run_plain(voss_ctx* ctx, uint64_t N, uint64_t seg_num,
                   voss_result* out);

/* Logical segments [seg_begin, seg_end) of a plain run, for callers that
 * run the range in pieces (checkpoint/resume, src/voss_v7_live.cu:599-621):
 * total_primes counts primes >= 7 of the range, first/last_prime are the
 * range's own (0 if none), and no head pairs are added (no (2,3), (3,5),
 * (5,7)).  Joining two consecutive pieces adds the pair (last of the
 * first, first of the second) to the main histogram only: pieces end on
 * SEG_NUM boundaries. */
int voss_run_plain_range(voss_ctx* ctx, uint64_t N, uint64_t seg_num,
                         uint64_t seg_begin, uint64_t seg_end, voss_result* out);

/* gap_sequence sample of a plain run (sample_gaps_kernel and its host logic,
 * src/voss_v7_live.cu:235-253, :526-547): gap_seq/pos_seq hold
 * VOSS_GAP_SEQ_SIZE entries each, zero where the reference writes nothing.
 * rows_out = number of rows with gap > 0. */
int voss_sample_plain(voss_ctx* ctx, uint64_t N, uint64_t seg_num,
                      int32_t* gap_seq, uint64_t* pos_seq, uint64_t* rows_out);

/* Interval [lo, hi) run.  No reference counterpart (the reference sieves
 * only [1, N]); it is the plain pipeline applied to an offset interval. */
int voss_run_interval(voss_ctx* ctx, uint64_t lo, uint64_t hi,
                      voss_result* out);

/* Many intervals [lo[i], hi[i]) in one persistent launch.  out[i]
 * receives interval i (out[i].large / large_capacity are honoured);
 * stats (optional) receives the batch's wall and device times in
 * seconds/device_ms/sieve_ms/windows (its other fields are zeroed). */
int voss_run_batch(voss_ctx* ctx, const uint64_t* lo, const uint64_t* hi,
                   size_t n, voss_interval_result* out, voss_result* stats);

/* The same intervals with one launch sequence per interval (the plain
 * per-interval form the batched form is measured against). */
int voss_run_batch_plain(voss_ctx* ctx, const uint64_t* lo,
                         const uint64_t* hi, size_t n,
                         voss_interval_result* out, voss_result* stats);

/* enable = 1 (the default): voss_run_batch captures its whole device
 * sequence (bucket fills, sieves, stitches of every round, read-back) in
 * one hipGraph and launches that.  enable = 0: ordinary stream launches,
 * which also time the bucket and sieve kernels separately (stats->sieve_ms
 * and stats->bucket_ms; under a graph they read -1). */
int voss_set_graph_mode(voss_ctx* ctx, int enable);

/* Prepared batch (no reference counterpart; for callers that sieve the
 * same interval set repeatedly).  voss_batch_prepare plans the windows,
 * rounds and bucket tasks once and gives the batch its own device buffers;
 * the first voss_batch_run captures the device sequence in a hipGraph and
 * every later run replays it.  The context's tuning must not change in
 * between (voss_batch_run then fails with VOSS_EINVAL).  Results are those
 * of voss_run_batch on the same intervals. */
typedef struct voss_batch voss_batch;
int voss_batch_prepare(voss_ctx* ctx, const uint64_t* lo, const uint64_t* hi,
                       size_t n, voss_batch** out);
int voss_batch_run(voss_batch* batch, voss_interval_result* out,
                   voss_result* stats);
/* Adds a Miller-Rabin batch over device-resident candidates (voss_device_alloc
 * buffers) to a prepared batch: every later voss_batch_run also sets
 * d_verdict[i] = 1 iff d_n[i] is prime (is_prime_u64, src/verify_cpp.cpp:38-57),
 * on a second stream beside the sieve rounds, inside the same captured
 * graph.  The buffers must stay valid until voss_batch_free. */
int voss_batch_add_mr(voss_batch* batch, const uint64_t* d_n, size_t count,
                      uint8_t* d_verdict);
void voss_batch_free(voss_batch* batch);

/* ----------------------------------------------------------- Miller-Rabin */

/* Deterministic 12-base Miller-Rabin on the GPU (64-bit Montgomery form):
 * replaces is_prime_u64 / mod_pow / mul_mod (src/verify_cpp.cpp:24-57).
 * verdict[i] = 1 iff n[i] is prime.  Host buffers. */
int voss_mr_batch(voss_ctx* ctx, const uint64_t* n, size_t count,
                  uint8_t* verdict);

/* Same on device-resident buffers (allocated with voss_device_alloc);
 * kernel_ms (optional) receives the device time of the MR kernels. */
int voss_mr_batch_device(voss_ctx* ctx, const uint64_t* d_n, size_t count,
                         uint8_t* d_verdict, double* kernel_ms);

/* No counterpart in the reference.  voss_mr_batch_device for callers that
 * cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc): the same kernels and
 * verdicts, with the survivor lists and counters in a fixed device array
 * of the library (32 MiB per GPU, kept until the process ends) instead of
 * context buffers that grow.  No hipMalloc, hipFree or other allocation.
 * count > VOSS_MR_NOALLOC_MAX returns VOSS_EINVAL.  Every call on one GPU
 * shares that array: calls that overlapped on one GPU would race (the
 * library admits one call at a time per process, and the call returns
 * after its kernels have finished). */
#define VOSS_MR_NOALLOC_MAX 4194304
int voss_mr_batch_device_noalloc(voss_ctx* ctx, const uint64_t* d_n, size_t count,
                                 uint8_t* d_verdict, double* kernel_ms);

/* verify_gap (src/verify_cpp.cpp:59-93) for a batch of gaps: both ends
 * prime (GPU Miller-Rabin) and every interior number composite (one GPU
 * wave per gap tests its interior with the same Miller-Rabin and stops at
 * the first prime).  Rows the reference treats specially get its verdicts:
 * gap < 2 verifies when both "ends" are prime; a gap larger than p_after
 * (p_before wraps past 0) never verifies.  ok[i] = 1 iff gap i verifies. */
int voss_verify_gaps(voss_ctx* ctx, const uint64_t* p_after,
                     const int32_t* gap, size_t count, uint8_t* ok);

/* ---------------------------------------------------------------- serving */

/* Continuous-batching server (no reference counterpart: the reference runs
 * one range per process).  A scheduler thread owns the context while the
 * server runs and admits queued requests, in arrival order, while request
 * slots and the HBM budget allow.  Two schedulers:
 *   VOSS_SERVE_DEVICE_LOOP (default)  a persistent kernel, one workgroup per
 *     CU, claims units of work (an interval's bucket tasks and LDS windows,
 *     chunks of a verify batch) from a ring in pinned host memory that the
 *     scheduler appends admitted requests to; a request's sums stay in its
 *     HBM slot, and the workgroup that completes its last unit joins the
 *     pairs across its windows and writes the result to the request's
 *     mailbox in pinned memory.  The scheduler never waits for the device
 *     while requests run.  The loop sieves 131072-byte windows (the
 *     context's tuning says so while the server runs).
 *   VOSS_SERVE_HOST_STEPS  each sieve step is up to windows_per_step LDS
 *     windows taken round-robin from all active interval requests, enqueued
 *     back to back and collected after their events; verify batches run on
 *     a second stream beside the steps.
 * Results are exactly those of voss_run_batch / voss_mr_batch on the same
 * request. */
typedef struct voss_server voss_server;

#define VOSS_SERVE_DEVICE_LOOP 0
#define VOSS_SERVE_HOST_STEPS 1
typedef struct {
    uint32_t max_active;       /* interval requests resident at once (default 64) */
    uint32_t windows_per_step; /* host steps: LDS windows per sieve round (default 4096) */
    uint64_t hbm_budget_bytes; /* admission budget for per-request device state
                                  (default 32 GiB) */
    uint32_t large_capacity;   /* large gaps kept per interval request (default 64) */
    uint32_t mr_per_step;      /* host steps: candidates per Miller-Rabin launch (default 2^22) */
    uint32_t mode;             /* VOSS_SERVE_DEVICE_LOOP or VOSS_SERVE_HOST_STEPS */
} voss_serve_config;

int voss_serve_start(voss_ctx* ctx, const voss_serve_config* cfg, voss_server** out);
int voss_serve_submit_interval(voss_server* s, uint64_t lo, uint64_t hi, uint64_t* id);
int voss_serve_submit_mr(voss_server* s, const uint64_t* n, size_t count, uint64_t* id);
/* *done = 1 when request id has finished; then out (interval requests, its
 * large buffer/capacity honoured) or verdict (MR requests, count bytes)
 * receive the result and latency_s the time from submit to completion. */
int voss_serve_poll(voss_server* s, uint64_t id, int* done, voss_interval_result* out,
                    uint8_t* verdict, double* latency_s);
int voss_serve_wait(voss_server* s, uint64_t id, voss_interval_result* out,
                    uint8_t* verdict, double* latency_s);
/* Counters of a running server (no reference counterpart): sieve steps
 * enqueued so far (device loop: launches of the loop kernel), LDS windows
 * enqueued so far (restarts included), the device time so far (host steps:
 * HIP events around each collected step; device loop: the workgroups' time
 * on units divided by the number of CUs, from the device wall clock), and
 * the wall time since voss_serve_start.  Any pointer may be NULL. */
int voss_serve_stats(voss_server* s, uint64_t* steps, uint64_t* windows, double* device_busy_s,
                     double* wall_s);
/* Device loop only (no reference counterpart): per unit kind (0 windows,
 * 1 bucket tasks, 2 Miller-Rabin chunks) the units run so far and the
 * workgroup-seconds they took.  units and busy_s hold 3 entries each (zeros
 * under the host-stepped scheduler). */
int voss_serve_unit_stats(voss_server* s, uint64_t* units, double* busy_s);
/* Finishes every submitted request, stops the scheduler, frees the server. */
int voss_serve_stop(voss_server* s);

/* --------------------------------------------------- device memory helpers */

int voss_device_alloc(voss_ctx* ctx, size_t bytes, void** dptr);
int voss_device_free(voss_ctx* ctx, void* dptr);
int voss_memcpy_h2d(voss_ctx* ctx, void* dst, const void* src, size_t bytes);
int voss_memcpy_d2h(voss_ctx* ctx, void* dst, const void* src, size_t bytes);

#ifdef __cplusplus
}
#endif
#endif
