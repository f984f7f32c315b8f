/*
 * voss_oracle.h -- CPU oracle for the VOSS prime-gap hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library, the
 * CLI binaries, the drop-in driver) links, loads or executes this code.  It
 * is called only by tests/, by __graft_entry__.smoke() as the checker, and by
 * bench.py's cpu_baseline leg.
 *
 * Every function restates the reference algorithm it names (file:line in
 * the reference tree, src/voss_v7_live.cu and src/verify_cpp.cpp).
 */
#ifndef VOSS_ORACLE_H
#define VOSS_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORACLE_MAX_GAP 100000      /* src/voss_v7_live.cu:23  MAX_GAP        */
#define ORACLE_MODQ_MAX_GAP 100    /* src/voss_v7_live.cu:25  MODQ_MAX_GAP   */
#define ORACLE_MODQ_ROWS 10        /* src/voss_v7_live.cu:28  MODQ_NUM_TOTAL */
#define ORACLE_Q_SLOTS 40          /* src/voss_v7_live.cu:196 MAX_Q_SLOTS    */
#define ORACLE_GAP_SEQ_SIZE 5000000 /* src/voss_v7_live.cu:35 GAP_SEQ_SIZE   */

typedef struct {
    uint64_t position; /* p_after (src/voss_v7_live.cu:295) */
    int32_t gap;
    float merit;       /* (float)((double)gap / log((double)p_before)) */
} oracle_large_gap;

typedef struct {
    /* counts; plain mode follows the reference conventions exactly:
     * total_primes starts at 3 (2,3,5 implicit), c1/c3 and modq count primes
     * >= 7 only (src/voss_v7_live.cu:404, :181-232). */
    uint64_t total_primes;
    uint64_t first_prime;   /* smallest prime counted (0 if none) */
    uint64_t last_prime;    /* largest prime counted (0 if none)  */
    uint64_t c1, c3;        /* primes >= 7 with p%4==1 / p%4==3   */
    uint64_t modq[ORACLE_Q_SLOTS]; /* q=6 at slots 0..5, q=30 at 6..35 */
    int64_t gaps[ORACLE_MAX_GAP];
    int64_t modq_gap[ORACLE_MODQ_ROWS * ORACLE_MODQ_MAX_GAP];
    uint64_t large_count;   /* untruncated number of large gaps */
    uint64_t sample_rows;   /* gap_sequence rows written (gap > 0) */
} oracle_result;

/* Number of primes 7 <= p <= limit written to out (at most cap).  Restates
 * generate_base_primes (src/voss_v7_live.cu:96-108).  Returns the total
 * count even when cap is smaller. */
size_t oracle_base_primes(uint64_t limit, uint32_t* out, size_t cap);

/* Plain [1,N] run with the reference's logical SEG_NUM segments
 * (src/voss_v7_live.cu:337-705).  large/large_cap: caller buffer for large
 * gaps in canonical order (merit desc, position asc).  gap_seq/pos_seq: if
 * non-NULL, arrays of ORACLE_GAP_SEQ_SIZE filled like the reference's
 * gap_seq_d/pos_seq_d (zero where not written).  threads<=0 means all.
 * Returns 0, or -1 on invalid arguments (SEG_NUM%30 != 0, N < 5). */
int oracle_run_plain(uint64_t N, uint64_t seg_num, int threads,
                     oracle_result* out, oracle_large_gap* large,
                     size_t large_cap, int32_t* gap_seq, uint64_t* pos_seq);

/* Interval [lo, hi) run: every prime lo <= p < hi (2, 3, 5 included when in
 * range), every consecutive pair inside the interval, no logical segments.
 * c1/c3/modq count every prime in the interval by its true residue. */
int oracle_run_interval(uint64_t lo, uint64_t hi, int threads,
                        oracle_result* out, oracle_large_gap* large,
                        size_t large_cap);

/* Deterministic 12-base Miller-Rabin, restating is_prime_u64
 * (src/verify_cpp.cpp:38-57) with mul_mod/mod_pow (:24-34). */
int oracle_is_prime_u64(uint64_t n);
void oracle_is_prime_batch(const uint64_t* n, size_t count, uint8_t* verdict,
                           int threads);

/* Modular multiplications the reference's mod_pow/is_prime_u64 performs on
 * n (the algorithmic work unit for MR): sum over bases tried of
 * bitlen(d)+popcount(d) plus the squarings the witness loop performs. */
uint64_t oracle_mr_mulmods(uint64_t n);

/* verify_gap (src/verify_cpp.cpp:59-93): both ends prime and every number
 * strictly between them composite.  Base primes are generated internally up
 * to isqrt(p_after)+1. */
int oracle_verify_gap(uint64_t p_after, int32_t gap);
void oracle_verify_gaps(const uint64_t* p_after, const int32_t* gap,
                        size_t count, uint8_t* ok, int threads);

/* merit as the reference computes it (src/voss_v7_live.cu:289-297) with the
 * host C library's log. */
float oracle_merit(uint64_t p_before, int32_t gap);

#ifdef __cplusplus
}
#endif
#endif
