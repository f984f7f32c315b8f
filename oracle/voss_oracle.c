/*
 * voss_oracle.c -- CPU oracle for the VOSS prime-gap hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library, never the thing
 * measured or shipped.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it.
 *
 * What it restates (reference tree paths):
 *   - wheel-30 bit layout: bit b of a segment <-> 30*(k_base+(b>>3)) +
 *     W30[b&7]   (src/voss_v7_live.cu:37-38, :189, :247-248, :556-557)
 *   - base primes 7..(int)sqrt((double)N)+1   (:96-108, :357)
 *   - crossing-off from max(p^2, seg_low)      (:449-478, :120-143)
 *   - logical segments [s*SEG_NUM+1, min((s+1)*SEG_NUM, N)]  (:434-446)
 *   - in-segment gaps: histogram, per-residue gap histogram for g<100,
 *     large-gap rule g>=500 or (g>100 and g/log(p_before)>=10)  (:255-312)
 *   - cross-segment gap added to the main histogram only     (:551-562)
 *   - mod-4 counts c1/c3 and q=6/q=30 counts of primes >= 7  (:181-232)
 *   - gap sample from the middle of each logical segment     (:526-547)
 *   - final +1 at g=1 and g=2 for (2,3) and (3,5)            (:634-635)
 *   - is_prime_u64 / mod_pow / mul_mod (src/verify_cpp.cpp:24-57)
 *   - verify_gap / build_base_sieve     (src/verify_cpp.cpp:59-103)
 *
 * The sieve itself is an ordinary segmented wheel-30 byte sieve over 1 MiB
 * chunks, parallel over chunks with OpenMP; chunks never straddle a logical
 * segment, so every in-chunk pair is an in-segment pair and only pairs that
 * join two chunks need the segment test.
 */
#include "voss_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static const int W30[8] = {1, 7, 11, 13, 17, 19, 23, 29};
static const int W30_TO_Q6_IDX[8] = {0, 0, 1, 0, 1, 0, 1, 1}; /* :30 */
static int8_t RES_IDX[30];
static int res_idx_ready = 0;

static void init_tables(void) {
    if (res_idx_ready) return;
    for (int i = 0; i < 30; i++) RES_IDX[i] = -1;
    for (int i = 0; i < 8; i++) RES_IDX[W30[i]] = (int8_t)i;
    res_idx_ready = 1;
}

#define CHUNK_BYTES (1u << 20) /* 1 MiB of wheel bytes = 31,457,280 numbers */

size_t oracle_base_primes(uint64_t limit, uint32_t* out, size_t cap) {
    if (limit < 7) return 0;
    uint8_t* s = (uint8_t*)malloc(limit + 1);
    memset(s, 1, limit + 1);
    s[0] = s[1] = 0;
    for (uint64_t i = 2; i * i <= limit; i++)
        if (s[i])
            for (uint64_t j = i * i; j <= limit; j += i) s[j] = 0;
    size_t cnt = 0;
    for (uint64_t i = 7; i <= limit; i++)
        if (s[i]) {
            if (out && cnt < cap) out[cnt] = (uint32_t)i;
            cnt++;
        }
    free(s);
    return cnt;
}

static uint64_t isqrt_u64(uint64_t n) {
    uint64_t r = (uint64_t)sqrtl((long double)n);
    while (r * r > n) r--;
    while ((r + 1) * (r + 1) <= n) r++;
    return r;
}

/* ---------------------------------------------------------------- sieve */

/* Sieve wheel bytes [k0,k1): buf[b] bit o = 1 iff 30*(k0+b)+W30[o] has no
 * prime factor p in bp with p*p <= that number.  Multiples are crossed from
 * max(p^2, 30*k0) upward, like the reference planner. */
static void sieve_bytes(uint64_t k0, uint64_t k1, const uint32_t* bp,
                        size_t nbp, uint8_t* buf) {
    size_t len = (size_t)(k1 - k0);
    memset(buf, 0xFF, len);
    uint64_t lo_num = 30 * k0, hi_num = 30 * k1;
    for (size_t i = 0; i < nbp; i++) {
        uint64_t p = bp[i];
        uint64_t p2 = p * p;
        if (p2 >= hi_num) break;
        uint64_t L = p2 > lo_num ? p2 : lo_num;
        uint64_t m0 = (L + p - 1) / p;
        uint64_t m0r = m0 % 30;
        for (int j = 0; j < 8; j++) {
            uint64_t m = m0 + (uint64_t)((W30[j] + 30 - (int)m0r) % 30);
            uint64_t n = p * m;
            uint64_t byte = n / 30;
            if (byte >= k1) continue;
            uint8_t mask = (uint8_t)~(1u << RES_IDX[n % 30]);
            for (uint64_t b = byte - k0; b < len; b += p) buf[b] &= mask;
        }
    }
    if (k0 == 0) buf[0] &= (uint8_t)~1u; /* the number 1 (:497-500) */
}

/* ------------------------------------------------------------ accumulate */

typedef struct {
    int64_t* gaps;     /* [ORACLE_MAX_GAP] */
    int64_t* modq_gap; /* [ORACLE_MODQ_ROWS*ORACLE_MODQ_MAX_GAP] */
    uint64_t c1, c3;
    uint64_t modq[ORACLE_Q_SLOTS];
    oracle_large_gap* lg;
    size_t lg_n, lg_cap;
} acc_t;

static void acc_init(acc_t* a) {
    memset(a, 0, sizeof(*a));
    a->gaps = (int64_t*)calloc(ORACLE_MAX_GAP, sizeof(int64_t));
    a->modq_gap = (int64_t*)calloc(ORACLE_MODQ_ROWS * ORACLE_MODQ_MAX_GAP,
                                   sizeof(int64_t));
}

static void acc_free(acc_t* a) {
    free(a->gaps);
    free(a->modq_gap);
    free(a->lg);
}

static void lg_push(acc_t* a, uint64_t pos, int32_t g, float m) {
    if (a->lg_n == a->lg_cap) {
        a->lg_cap = a->lg_cap ? 2 * a->lg_cap : 1024;
        a->lg = (oracle_large_gap*)realloc(a->lg,
                                           a->lg_cap * sizeof(oracle_large_gap));
    }
    a->lg[a->lg_n].position = pos;
    a->lg[a->lg_n].gap = g;
    a->lg[a->lg_n].merit = m;
    a->lg_n++;
}

float oracle_merit(uint64_t p_before, int32_t gap) {
    return (float)((double)gap / log((double)p_before));
}

/* One consecutive pair (p, q).  full=0: main histogram only (a pair that
 * the reference stitches on the host, :559-562). */
static inline void pair_acc(acc_t* a, uint64_t p, uint64_t q, int full) {
    uint64_t g = q - p;
    if (g == 0) return;
    if (g < ORACLE_MAX_GAP) a->gaps[g]++;
    if (!full || p < 7) return;
    if (g < ORACLE_MODQ_MAX_GAP) {
        int w = RES_IDX[p % 30];
        a->modq_gap[W30_TO_Q6_IDX[w] * ORACLE_MODQ_MAX_GAP + g]++;
        a->modq_gap[(2 + w) * ORACLE_MODQ_MAX_GAP + g]++;
    }
    double merit_d = 0.0;
    if (g > 100) merit_d = (double)g / log((double)p);
    if (g >= 500 || merit_d >= 10.0) {
        if (g < ORACLE_MAX_GAP) lg_push(a, q, (int32_t)g, (float)merit_d);
    }
}

static inline void count_prime(acc_t* a, uint64_t v) {
    uint32_t r4 = (uint32_t)(v & 3);
    if (r4 == 1) a->c1++;
    else if (r4 == 3) a->c3++;
    a->modq[v % 6]++;
    a->modq[6 + v % 30]++;
}

/* ----------------------------------------------------------------- chunks */

typedef struct {
    uint64_t k0, k1;  /* wheel bytes [k0,k1) */
    uint64_t lo, hi;  /* inclusive number bounds */
    int64_t seg;      /* logical segment id */
    uint64_t count, first, last;
} chunk_t;

/* Walks the primes of a sieved chunk in increasing order. */
#define FOR_EACH_PRIME(buf, c, v, BODY)                                      \
    do {                                                                     \
        size_t _len = (size_t)((c)->k1 - (c)->k0);                           \
        for (size_t _b = 0; _b < _len; _b += 8) {                            \
            uint64_t _w = 0;                                                 \
            size_t _n = _len - _b < 8 ? _len - _b : 8;                       \
            memcpy(&_w, (buf) + _b, _n);                                     \
            while (_w) {                                                     \
                int _t = __builtin_ctzll(_w);                                \
                _w &= _w - 1;                                                \
                uint64_t v = 30 * ((c)->k0 + _b + (uint64_t)(_t >> 3)) +     \
                             (uint64_t)W30[_t & 7];                          \
                if (v < (c)->lo || v > (c)->hi) continue;                    \
                BODY                                                         \
            }                                                                \
        }                                                                    \
    } while (0)

static void process_chunk(chunk_t* c, const uint32_t* bp, size_t nbp,
                          uint8_t* buf, acc_t* a) {
    sieve_bytes(c->k0, c->k1, bp, nbp, buf);
    uint64_t prev = 0, cnt = 0, first = 0;
    FOR_EACH_PRIME(buf, c, v, {
        count_prime(a, v);
        if (cnt == 0) first = v;
        else pair_acc(a, prev, v, 1);
        prev = v;
        cnt++;
    });
    c->count = cnt;
    c->first = first;
    c->last = prev;
}

static int lg_cmp(const void* x, const void* y) {
    const oracle_large_gap* a = (const oracle_large_gap*)x;
    const oracle_large_gap* b = (const oracle_large_gap*)y;
    if (a->merit != b->merit) return a->merit > b->merit ? -1 : 1;
    if (a->position != b->position) return a->position < b->position ? -1 : 1;
    return 0;
}

/* Append the chunks covering numbers [a,b] (logical segment seg). */
static void add_chunks(chunk_t** ch, size_t* n, size_t* cap, uint64_t a,
                       uint64_t b, int64_t seg) {
    uint64_t kb0 = a / 30, kb1 = b / 30 + 1;
    for (uint64_t k = kb0; k < kb1; k += CHUNK_BYTES) {
        if (*n == *cap) {
            *cap = *cap ? 2 * *cap : 1024;
            *ch = (chunk_t*)realloc(*ch, *cap * sizeof(chunk_t));
        }
        chunk_t* c = &(*ch)[(*n)++];
        memset(c, 0, sizeof(*c));
        c->k0 = k;
        c->k1 = (k + CHUNK_BYTES < kb1) ? k + CHUNK_BYTES : kb1;
        c->lo = a;
        c->hi = b;
        c->seg = seg;
    }
}

static void run_chunks(chunk_t* ch, size_t nch, const uint32_t* bp, size_t nbp,
                       int threads, acc_t* total) {
    int nt = 1;
#ifdef _OPENMP
    nt = threads > 0 ? threads : omp_get_max_threads();
#endif
    acc_t* accs = (acc_t*)calloc((size_t)nt, sizeof(acc_t));
    for (int t = 0; t < nt; t++) acc_init(&accs[t]);
#ifdef _OPENMP
#pragma omp parallel num_threads(nt)
#endif
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        uint8_t* buf = (uint8_t*)malloc(CHUNK_BYTES + 8);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
        for (long i = 0; i < (long)nch; i++)
            process_chunk(&ch[i], bp, nbp, buf, &accs[tid]);
        free(buf);
    }
    for (int t = 0; t < nt; t++) {
        acc_t* a = &accs[t];
        for (int g = 0; g < ORACLE_MAX_GAP; g++) total->gaps[g] += a->gaps[g];
        for (int i = 0; i < ORACLE_MODQ_ROWS * ORACLE_MODQ_MAX_GAP; i++)
            total->modq_gap[i] += a->modq_gap[i];
        total->c1 += a->c1;
        total->c3 += a->c3;
        for (int i = 0; i < ORACLE_Q_SLOTS; i++) total->modq[i] += a->modq[i];
        for (size_t i = 0; i < a->lg_n; i++)
            lg_push(total, a->lg[i].position, a->lg[i].gap, a->lg[i].merit);
        acc_free(a);
    }
    free(accs);
}

static void finish_result(oracle_result* out, acc_t* acc,
                          oracle_large_gap* large, size_t large_cap) {
    memcpy(out->gaps, acc->gaps, sizeof(out->gaps));
    memcpy(out->modq_gap, acc->modq_gap, sizeof(out->modq_gap));
    out->c1 = acc->c1;
    out->c3 = acc->c3;
    memcpy(out->modq, acc->modq, sizeof(out->modq));
    out->large_count = acc->lg_n;
    qsort(acc->lg, acc->lg_n, sizeof(oracle_large_gap), lg_cmp);
    if (large) {
        size_t n = acc->lg_n < large_cap ? acc->lg_n : large_cap;
        memcpy(large, acc->lg, n * sizeof(oracle_large_gap));
    }
}

/* Primes of logical segment s with in-segment index in [i0, i0+cnt) ->
 * vals.  Re-sieves only the chunks that hold them. */
static void segment_primes(chunk_t* ch, size_t c_begin, size_t c_end,
                           uint64_t i0, uint64_t cnt, const uint32_t* bp,
                           size_t nbp, uint64_t* vals) {
    uint8_t* buf = (uint8_t*)malloc(CHUNK_BYTES + 8);
    uint64_t idx = 0, got = 0;
    for (size_t c = c_begin; c < c_end && got < cnt; c++) {
        if (idx + ch[c].count <= i0) {
            idx += ch[c].count;
            continue;
        }
        sieve_bytes(ch[c].k0, ch[c].k1, bp, nbp, buf);
        FOR_EACH_PRIME(buf, &ch[c], v, {
            if (idx >= i0 && got < cnt) vals[got++] = v;
            idx++;
        });
    }
    free(buf);
}

int oracle_run_plain(uint64_t N, uint64_t seg_num, int threads,
                     oracle_result* out, oracle_large_gap* large,
                     size_t large_cap, int32_t* gap_seq, uint64_t* pos_seq) {
    if (!out || N < 5 || seg_num == 0 || seg_num % 30 != 0) return -1;
    init_tables();
    memset(out, 0, sizeof(*out));
    uint64_t limit = (uint64_t)((int)sqrt((double)N) + 1); /* :357 */
    size_t nbp = oracle_base_primes(limit, NULL, 0);
    uint32_t* bp = (uint32_t*)malloc((nbp + 1) * sizeof(uint32_t));
    oracle_base_primes(limit, bp, nbp);

    uint64_t num_seg = (N + seg_num - 1) / seg_num; /* :342 */
    chunk_t* ch = NULL;
    size_t nch = 0, cap = 0;
    size_t* seg_first_chunk = (size_t*)malloc((num_seg + 1) * sizeof(size_t));
    for (uint64_t s = 0; s < num_seg; s++) {
        uint64_t a = s * seg_num + 1;
        uint64_t b = (s + 1) * seg_num;
        if (b > N) b = N;
        seg_first_chunk[s] = nch;
        add_chunks(&ch, &nch, &cap, a, b, (int64_t)s);
    }
    seg_first_chunk[num_seg] = nch;

    acc_t acc;
    acc_init(&acc);
    run_chunks(ch, nch, bp, nbp, threads, &acc);

    /* stitch chunk edges; the implicit prime 5 belongs to no segment */
    uint64_t last = 5, total = 3;
    int64_t last_seg = -1;
    for (size_t i = 0; i < nch; i++) {
        if (ch[i].count == 0) continue;
        pair_acc(&acc, last, ch[i].first, ch[i].seg == last_seg);
        last = ch[i].last;
        last_seg = ch[i].seg;
        total += ch[i].count;
    }
    acc.gaps[1] += 1; /* (2,3)  :634 */
    acc.gaps[2] += 1; /* (3,5)  :635 */

    finish_result(out, &acc, large, large_cap);
    out->total_primes = total;
    out->first_prime = 2;
    out->last_prime = last;

    /* gap_sequence sample (:526-547) */
    if (gap_seq && pos_seq) {
        memset(gap_seq, 0, ORACLE_GAP_SEQ_SIZE * sizeof(int32_t));
        memset(pos_seq, 0, ORACLE_GAP_SEQ_SIZE * sizeof(uint64_t));
        int64_t per_seg = ORACLE_GAP_SEQ_SIZE / (int64_t)num_seg;
        uint64_t rows = 0;
        for (uint64_t s = 0; s < num_seg && per_seg > 0; s++) {
            uint64_t n_pos = 0;
            for (size_t c = seg_first_chunk[s]; c < seg_first_chunk[s + 1]; c++)
                n_pos += ch[c].count;
            if (n_pos <= 1) continue;
            int64_t start = (int64_t)(n_pos / 2);
            int64_t avail = (int64_t)n_pos - start - 1;
            int64_t sample_n = avail < per_seg ? avail : per_seg;
            if (sample_n <= 1) continue;
            int64_t offset = (int64_t)s * per_seg;
            if (offset + sample_n > ORACLE_GAP_SEQ_SIZE) continue;
            uint64_t* vals =
                (uint64_t*)malloc((size_t)(sample_n + 1) * sizeof(uint64_t));
            segment_primes(ch, seg_first_chunk[s], seg_first_chunk[s + 1],
                           (uint64_t)start, (uint64_t)sample_n + 1, bp, nbp,
                           vals);
            for (int64_t i = 0; i < sample_n; i++) {
                gap_seq[offset + i] = (int32_t)(vals[i + 1] - vals[i]);
                pos_seq[offset + i] = vals[i];
                rows++;
            }
            free(vals);
        }
        out->sample_rows = rows;
    }

    acc_free(&acc);
    free(seg_first_chunk);
    free(ch);
    free(bp);
    return 0;
}

int oracle_run_interval(uint64_t lo, uint64_t hi, int threads,
                        oracle_result* out, oracle_large_gap* large,
                        size_t large_cap) {
    if (!out || hi < lo) return -1;
    init_tables();
    memset(out, 0, sizeof(*out));
    acc_t acc;
    acc_init(&acc);
    uint64_t total = 0, first = 0, last = 0;
    /* 2, 3, 5 are not on the wheel: they form a prefix of the interval */
    static const uint64_t small[3] = {2, 3, 5};
    for (int i = 0; i < 3; i++) {
        uint64_t p = small[i];
        if (p < lo || p >= hi) continue;
        count_prime(&acc, p);
        if (total == 0) first = p;
        else pair_acc(&acc, last, p, 1);
        last = p;
        total++;
    }
    if (hi > 7 && hi - 1 >= (lo > 7 ? lo : 7)) {
        uint64_t a = lo > 7 ? lo : 7, b = hi - 1;
        uint64_t limit = isqrt_u64(b) + 1;
        size_t nbp = oracle_base_primes(limit, NULL, 0);
        uint32_t* bp = (uint32_t*)malloc((nbp + 1) * sizeof(uint32_t));
        oracle_base_primes(limit, bp, nbp);
        chunk_t* ch = NULL;
        size_t nch = 0, cap = 0;
        add_chunks(&ch, &nch, &cap, a, b, 0);
        run_chunks(ch, nch, bp, nbp, threads, &acc);
        for (size_t i = 0; i < nch; i++) {
            if (ch[i].count == 0) continue;
            if (total == 0) first = ch[i].first;
            else pair_acc(&acc, last, ch[i].first, 1);
            last = ch[i].last;
            total += ch[i].count;
        }
        free(ch);
        free(bp);
    }
    finish_result(out, &acc, large, large_cap);
    out->total_primes = total;
    out->first_prime = first;
    out->last_prime = last;
    acc_free(&acc);
    return 0;
}

/* ----------------------------------------------------------- Miller-Rabin */

static inline uint64_t mul_mod(uint64_t a, uint64_t b, uint64_t m) {
    return (uint64_t)((unsigned __int128)a * b % m);
}

static inline uint64_t mod_pow(uint64_t base, uint64_t e, uint64_t mod) {
    uint64_t result = 1;
    base %= mod;
    while (e > 0) {
        if (e & 1) result = mul_mod(result, base, mod);
        base = mul_mod(base, base, mod);
        e >>= 1;
    }
    return result;
}

static const uint64_t BASES[12] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};

int oracle_is_prime_u64(uint64_t n) {
    if (n < 2) return 0;
    for (int i = 0; i < 12; i++) {
        if (n == BASES[i]) return 1;
        if (n % BASES[i] == 0) return 0;
    }
    uint64_t d = n - 1;
    int r = 0;
    while ((d & 1) == 0) {
        d >>= 1;
        r++;
    }
    for (int i = 0; i < 12; i++) {
        uint64_t x = mod_pow(BASES[i], d, n);
        if (x == 1 || x == n - 1) continue;
        int composite = 1;
        for (int j = 0; j < r - 1; j++) {
            x = mul_mod(x, x, n);
            if (x == n - 1) {
                composite = 0;
                break;
            }
        }
        if (composite) return 0;
    }
    return 1;
}

void oracle_is_prime_batch(const uint64_t* n, size_t count, uint8_t* verdict,
                           int threads) {
#ifdef _OPENMP
    int nt = threads > 0 ? threads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 4096) num_threads(nt)
#endif
    for (long i = 0; i < (long)count; i++)
        verdict[i] = (uint8_t)oracle_is_prime_u64(n[i]);
    (void)threads;
}

uint64_t oracle_mr_mulmods(uint64_t n) {
    if (n < 2) return 0;
    for (int i = 0; i < 12; i++)
        if (n == BASES[i] || n % BASES[i] == 0) return 0;
    uint64_t d = n - 1;
    int r = 0;
    while ((d & 1) == 0) {
        d >>= 1;
        r++;
    }
    uint64_t per_pow = (uint64_t)(64 - __builtin_clzll(d)) +
                       (uint64_t)__builtin_popcountll(d);
    uint64_t work = 0;
    for (int i = 0; i < 12; i++) {
        uint64_t x = mod_pow(BASES[i], d, n);
        work += per_pow;
        if (x == 1 || x == n - 1) continue;
        int composite = 1;
        for (int j = 0; j < r - 1; j++) {
            x = mul_mod(x, x, n);
            work++;
            if (x == n - 1) {
                composite = 0;
                break;
            }
        }
        if (composite) return work;
    }
    return work;
}

/* ---------------------------------------------------------------- verify */

static int verify_gap_bp(uint64_t p_after, int32_t gap, const uint32_t* bp,
                         size_t nbp) {
    uint64_t p_before = p_after - (uint64_t)gap;
    if (!oracle_is_prime_u64(p_before)) return 0;
    if (!oracle_is_prime_u64(p_after)) return 0;
    if (gap < 2) return 1;
    int n_int = gap - 1;
    uint8_t* is_comp = (uint8_t*)calloc((size_t)n_int, 1);
    for (int i = 0; i < n_int; i += 2) is_comp[i] = 1;
    uint64_t sq = isqrt_u64(p_after);
    for (size_t k = 0; k < nbp; k++) {
        uint64_t q = bp[k];
        if (q == 2) continue;
        if (q > sq) break;
        uint64_t start = (p_before / q + 1) * q;
        if ((start & 1) == 0) start += q;
        if (start >= p_after) continue;
        int64_t idx = (int64_t)(start - p_before - 1);
        if (idx < 0) idx = 0;
        int64_t step = (int64_t)(2 * q);
        for (int64_t i = idx; i < n_int; i += step) is_comp[i] = 1;
    }
    int ok = 1;
    for (int i = 0; i < n_int; i++)
        if (!is_comp[i]) {
            ok = 0;
            break;
        }
    free(is_comp);
    return ok;
}

/* all primes 2..limit, like build_base_sieve (src/verify_cpp.cpp:95-103) */
static uint32_t* all_primes(uint64_t limit, size_t* n) {
    uint8_t* s = (uint8_t*)malloc(limit + 1);
    memset(s, 1, limit + 1);
    s[0] = 0;
    if (limit >= 1) s[1] = 0;
    for (uint64_t i = 2; i * i <= limit; i++)
        if (s[i])
            for (uint64_t j = i * i; j <= limit; j += i) s[j] = 0;
    size_t cnt = 0;
    for (uint64_t i = 2; i <= limit; i++) cnt += s[i];
    uint32_t* out = (uint32_t*)malloc((cnt + 1) * sizeof(uint32_t));
    cnt = 0;
    for (uint64_t i = 2; i <= limit; i++)
        if (s[i]) out[cnt++] = (uint32_t)i;
    free(s);
    *n = cnt;
    return out;
}

int oracle_verify_gap(uint64_t p_after, int32_t gap) {
    size_t nbp = 0;
    uint32_t* bp = all_primes(isqrt_u64(p_after) + 1, &nbp);
    int ok = verify_gap_bp(p_after, gap, bp, nbp);
    free(bp);
    return ok;
}

void oracle_verify_gaps(const uint64_t* p_after, const int32_t* gap,
                        size_t count, uint8_t* ok, int threads) {
    uint64_t max_p = 0;
    for (size_t i = 0; i < count; i++)
        if (p_after[i] > max_p) max_p = p_after[i];
    size_t nbp = 0;
    uint32_t* bp = all_primes(isqrt_u64(max_p) + 1, &nbp);
#ifdef _OPENMP
    int nt = threads > 0 ? threads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 64) num_threads(nt)
#endif
    for (long i = 0; i < (long)count; i++)
        ok[i] = (uint8_t)verify_gap_bp(p_after[i], gap[i], bp, nbp);
    (void)threads;
    free(bp);
}
