// api.hip -- the C ABI of include/voss/voss.h: context, planning of LDS
// windows, kernel launches and the host finalisation that turns the device
// accumulators into the reference's output arrays.  Serving is in serve.hip.
#include "engine.h"

namespace voss {
namespace core {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

int hip_fail(hipError_t e, const char* what) {
    return fail(VOSS_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

uint64_t isqrt_u64(uint64_t n) {
    uint64_t r = (uint64_t)std::sqrt((long double)n);
    while (r > 0 && r * r > n) r--;
    while ((r + 1) * (r + 1) <= n) r++;
    return r;
}

// Primes 7 <= p <= limit (segmented odd-only sieve; limit < 2^32).
std::vector<uint32_t> make_base_primes(uint64_t limit) {
    std::vector<uint32_t> out;
    if (limit < 7) return out;
    const uint64_t root = isqrt_u64(limit);
    std::vector<uint8_t> small(root + 1, 1);
    std::vector<uint32_t> sp;
    for (uint64_t i = 2; i <= root; i++) {
        if (!small[i]) continue;
        if (i > 2) sp.push_back((uint32_t)i);
        for (uint64_t j = i * i; j <= root; j += i) small[j] = 0;
    }
    const uint64_t SEG = 1 << 20;  // odd numbers per segment
    std::vector<uint8_t> seg(SEG);
    for (uint64_t lo = 1; lo <= limit; lo += 2 * SEG) {
        // seg[i] <-> lo + 2i
        const uint64_t hi = std::min<uint64_t>(lo + 2 * SEG - 1, limit);
        const uint64_t n = (hi - lo) / 2 + 1;
        std::fill(seg.begin(), seg.begin() + n, 1);
        for (uint32_t q : sp) {
            const uint64_t q2 = (uint64_t)q * q;
            if (q2 > hi) break;
            uint64_t st = q2 >= lo ? q2 : ((lo + q - 1) / q) * q;
            if ((st & 1) == 0) st += q;
            for (uint64_t m = (st - lo) / 2; m < n; m += q) seg[m] = 0;
        }
        for (uint64_t i = 0; i < n; i++) {
            const uint64_t v = lo + 2 * i;
            if (seg[i] && v >= 7) out.push_back((uint32_t)v);
        }
    }
    return out;
}

float merit_of(uint64_t p_before, uint32_t g) {
    return (float)((double)g / std::log((double)p_before));
}

bool merit_ge10(uint64_t p_before, uint32_t g) {
    return (double)g / std::log((double)p_before) >= 10.0;
}

// memcpy on several host threads for large blocks (pinned staging of the
// window descriptors: 1.2 GB for a plain run to 1e14).
void par_copy(void* dst, const void* src, size_t bytes) {
    const size_t kChunk = 1u << 22;
    parallel_for((bytes + kChunk - 1) / kChunk, [&](size_t c0, size_t c1) {
        const size_t b0 = c0 * kChunk, b1 = std::min(bytes, c1 * kChunk);
        std::memcpy(static_cast<char*>(dst) + b0, static_cast<const char*>(src) + b0, b1 - b0);
    }, 2);
}

// The process-wide gate of library calls (call_gate.h), and whether the
// calling thread is a server's scheduler.
CallGate& gate() {
    static CallGate g;
    return g;
}

thread_local bool t_scheduler = false;

int set_device(voss_ctx* c) {
    HIPCHK(hipSetDevice(c->device));
    return VOSS_OK;
}

// Base primes covering every window that ends at or below max_number.
int ensure_primes(voss_ctx* c, uint64_t max_number) {
    const uint64_t need = isqrt_u64(max_number) + 1;
    if (need <= c->primes_limit && c->d_nprimes == c->primes.size()) return VOSS_OK;
    if (need > 0xFFFFFFFFull) return fail(VOSS_EINVAL, "range beyond 2^64");
    uint64_t lim = std::max<uint64_t>(need, std::max<uint64_t>(c->primes_limit * 2, 1 << 16));
    lim = std::min<uint64_t>(lim, 0xFFFFFFFFull);
    c->primes = make_base_primes(lim);
    c->primes_limit = lim;
    c->inv_prefix.assign(c->primes.size() + 1, 0.0);
    for (size_t i = 0; i < c->primes.size(); i++)
        c->inv_prefix[i + 1] = c->inv_prefix[i] + 1.0 / c->primes[i];
    std::vector<PrimeRec> rec(c->primes.size());
    int residx[30];
    for (int i = 0; i < 30; i++) residx[i] = -1;
    for (int o = 0; o < 8; o++) residx[kW30[o]] = o;
    for (size_t i = 0; i < c->primes.size(); i++) {
        const uint32_t p = c->primes[i];
        uint32_t w = 0;
        for (int j = 0; j < 8; j++)
            w |= (uint32_t)residx[(uint64_t)p * (uint64_t)kW30[j] % 30] << (3 * j);
        rec[i].p = p;
        rec[i].perm = w | (p % 30u) << 24;
        rec[i].magic = ~0ull / p;
    }
    const size_t n = std::max<size_t>(c->primes.size(), 1);
    HIPCHK(c->d_primes.reserve(n));
    HIPCHK(c->d_p32.reserve(n));
    if (!c->primes.empty()) {
        HIPCHK(hipMemcpy(c->d_primes.p, rec.data(), rec.size() * sizeof(PrimeRec),
                         hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(c->d_p32.p, c->primes.data(), c->primes.size() * sizeof(uint32_t),
                         hipMemcpyHostToDevice));
    }
    c->d_nprimes = c->primes.size();
    return VOSS_OK;
}

uint32_t primes_upto(const voss_ctx* c, uint64_t max_number) {
    const uint64_t r = isqrt_u64(max_number);
    return (uint32_t)(std::upper_bound(c->primes.begin(), c->primes.end(),
                                       (uint32_t)std::min<uint64_t>(r, 0xFFFFFFFFull)) -
                      c->primes.begin());
}

// Presieve tables: for each group of three primes, one word per residue
// of the word index modulo the product (a period of the group's pattern).
int build_patterns(voss_ctx* c) {
    std::vector<uint32_t> all;
    for (int g = 0; g < kPresieveGroups; g++) {
        const int* q = kPresieveGroupPrimes[g];
        const uint32_t P = (uint32_t)(q[0] * q[1] * q[2]);
        c->pat_off[g] = (uint32_t)all.size();
        c->pat_len[g] = P;
        std::vector<uint8_t> bytes(4 * (size_t)P, 0xFF);
        for (int k = 0; k < 3; k++) {
            const uint32_t p = (uint32_t)q[k];
            if (p == 1) continue;
            // crossing every multiple of p (the prime itself is restored in
            // window 0)
            for (int o = 0; o < 8; o++) {
                // first byte b with 30b + W30[o] = 0 mod p
                uint32_t b0 = 0;
                while ((30u * b0 + kW30Mask[o]) % p) b0++;
                for (uint64_t b = b0; b < bytes.size(); b += p) bytes[b] &= (uint8_t)~(1u << o);
            }
        }
        for (uint32_t w = 0; w <= P; w++) {  // word P repeats word 0
            uint32_t v;
            std::memcpy(&v, &bytes[4 * (size_t)(w % P)], 4);
            all.push_back(v);
        }
    }
    if (all.size() != (size_t)kPresieveWordsPadded) return fail(VOSS_EDEVICE, "presieve table size");
    // word 0 (numbers 1..119): put back the presieved primes themselves
    uint32_t restore = 0;
    for (int g = 0; g < kPresieveGroups; g++)
        for (int k = 0; k < 3; k++) {
            const uint32_t p = (uint32_t)kPresieveGroupPrimes[g][k];
            if (p == 1) continue;
            int o = -1;
            for (int i = 0; i < 8; i++) if ((uint32_t)kW30[i] == p % 30) o = i;
            restore |= 1u << ((p / 30) * 8 + o);
        }
    c->restore0 = restore;
    HIPCHK(c->d_pat.reserve(all.size()));
    HIPCHK(hipMemcpy(c->d_pat.p, all.data(), all.size() * 4, hipMemcpyHostToDevice));
    return VOSS_OK;
}

// Largest p_before for which float merit would reach 10 at gap g:
// (double)g / log((double)p) >= 10.0 is monotone in p, so binary search.
int build_lg_thresh(voss_ctx* c) {
    c->lg_thresh.assign(kHistBins, 0);
    for (int h = 0; h < kHistBins; h++) {
        const uint32_t g = 2u * (uint32_t)h;
        if (g <= 100) { c->lg_thresh[h] = 0; continue; }
        if (g >= (uint32_t)kLargeAlways) { c->lg_thresh[h] = ~0ull; continue; }
        if (merit_ge10(~0ull, g)) { c->lg_thresh[h] = ~0ull; continue; }
        uint64_t lo = 2, hi = ~0ull;  // pred(lo) true, pred(hi) false
        while (hi - lo > 1) {
            const uint64_t mid = lo + (hi - lo) / 2;
            if (merit_ge10(mid, g)) lo = mid; else hi = mid;
        }
        c->lg_thresh[h] = lo;
    }
    HIPCHK(c->d_lg_thresh.reserve(kHistBins));
    HIPCHK(hipMemcpy(c->d_lg_thresh.p, c->lg_thresh.data(), kHistBins * 8,
                     hipMemcpyHostToDevice));
    return VOSS_OK;
}

void host_pair(JobHost& J, uint64_t p, uint64_t q, bool full) {
    J.init();
    const uint64_t g = q - p;
    if (g == 0) return;
    if (g == 1) {
        J.gap1++;
        return;  // (2,3): no residue row, never large
    }
    if (g >= 2ull * kHistBins || (g & 1)) {
        J.overflow = true;
        return;
    }
    J.hist2[g / 2]++;
    if (!full || p < 7) return;
    if (g < VOSS_MODQ_MAX_GAP) {
        static const int q6[8] = {0, 0, 1, 0, 1, 0, 1, 1};
        int o = -1;
        for (int i = 0; i < 8; i++) if ((uint64_t)kW30[i] == p % 30) o = i;
        J.mq2[q6[o] * kModqBins + g / 2]++;
        J.mq2[(2 + o) * kModqBins + g / 2]++;
    }
    if (g > 100 && (g >= (uint64_t)kLargeAlways || merit_ge10(p, (uint32_t)g))) {
        J.large_count++;
        J.large.push_back({q, (int32_t)g, merit_of(p, (uint32_t)g)});
    }
}

void add_acc(JobHost& J, const JobAcc& A) {
    J.init();
    for (int h = 1; h < kHistBins; h++) J.hist2[h] += (int64_t)A.hist[h];
    for (int i = 0; i < kModqRows * kModqBins; i++)
        if (i % kModqBins) J.mq2[i] += (int64_t)A.modq[i];
    J.count += A.count;
    J.c1 += A.c1;
    for (int o = 0; o < 8; o++) J.mod30[o] += A.mod30[o];
    if (A.count) {
        if (J.first == 0 || A.first < J.first) J.first = A.first;
        if (A.last > J.last) J.last = A.last;
    }
    J.large_count += A.large_count;
}

// Canonical order of the large-gap list: merit descending, then position
// ascending (the reference sorts by merit alone with an unstable sort,
// src/voss_v7_live.cu:813, so ties have no defined order there).  Merits
// are positive floats, whose bit patterns order like the values: an LSD
// radix sort (11-bit digits, digits that never vary skipped) of the
// inverted merit bits carrying the list index, then runs of equal merit
// sorted by position.
void sort_large(LargeVec& v) {
    const size_t n = v.size();
    if (n < 2) return;
    // scratch kept per thread for lists up to 2^20 records (fresh multi-MB
    // vectors per call cost page faults), local beyond
    static thread_local std::vector<uint32_t> keep_u32[4];
    static thread_local LargeVec keep_out;
    std::vector<uint32_t> own_u32[4];
    LargeVec own_out;
    const bool keep = n <= (1u << 20);
    std::vector<uint32_t>* u = keep ? keep_u32 : own_u32;
    LargeVec& out = keep ? keep_out : own_out;
    for (int q = 0; q < 4; q++) u[q].resize(n);
    uint32_t *key = u[0].data(), *idx = u[1].data(), *key2 = u[2].data(), *idx2 = u[3].data();
    uint32_t all_or = 0, all_and = ~0u;
    for (size_t i = 0; i < n; i++) {
        uint32_t mb;
        std::memcpy(&mb, &v[i].merit, 4);
        key[i] = 0xFFFFFFFFu - mb;
        idx[i] = (uint32_t)i;
        all_or |= key[i];
        all_and &= key[i];
    }
    for (int sh = 0; sh < 32; sh += 11) {
        const uint32_t dmask = 0x7FFu;
        if ((((all_or ^ all_and) >> sh) & dmask) == 0) continue;  // digit constant
        size_t cnt[2049] = {};
        for (size_t i = 0; i < n; i++) cnt[((key[i] >> sh) & dmask) + 1]++;
        for (int d = 0; d < 2048; d++) cnt[d + 1] += cnt[d];
        for (size_t i = 0; i < n; i++) {
            const size_t at = cnt[(key[i] >> sh) & dmask]++;
            key2[at] = key[i];
            idx2[at] = idx[i];
        }
        std::swap(key, key2);
        std::swap(idx, idx2);
    }
    out.resize(n);
    for (size_t i = 0; i < n; i++) out[i] = v[idx[i]];
    for (size_t i = 0; i < n;) {
        size_t j = i + 1;
        while (j < n && key[j] == key[i]) j++;
        if (j - i > 1)
            std::sort(out.begin() + i, out.begin() + j,
                      [](const voss_large_gap& x, const voss_large_gap& y) {
                          return x.position < y.position;
                      });
        i = j;
    }
    v.swap(out);
}

// The canonical order of sort_large as a comparison.
bool large_before(const voss_large_gap& x, const voss_large_gap& y) {
    return x.merit > y.merit || (x.merit == y.merit && x.position < y.position);
}

// The k first records of the canonical order into dst (k <= v.size()):
// each host thread selects its chunk's best k, then the best k of those.
// For long lists of which the caller keeps a few (a plain run to 1e14 has
// 4.2e7 candidates).
void top_large(const LargeVec& v, size_t k, voss_large_gap* dst) {
    std::mutex m;
    std::vector<voss_large_gap> best;
    parallel_for(v.size(), [&](size_t i0, size_t i1) {
        std::vector<voss_large_gap> part(std::min(k, i1 - i0));
        std::partial_sort_copy(v.begin() + i0, v.begin() + i1, part.begin(), part.end(),
                               large_before);
        std::lock_guard<std::mutex> g(m);
        best.insert(best.end(), part.begin(), part.end());
    }, 1u << 16);
    std::partial_sort_copy(best.begin(), best.end(), dst, dst + k, large_before);
}

// Histograms and large gaps into the reference's array layout.
void write_result(JobHost& J, voss_result* out) {
    J.init();
    std::memset(out->gaps, 0, sizeof(out->gaps));
    std::memset(out->modq_gap, 0, sizeof(out->modq_gap));
    out->gaps[1] = J.gap1;
    for (int h = 1; h < kHistBins; h++) out->gaps[2 * h] = J.hist2[h];
    for (int r = 0; r < kModqRows; r++)
        for (int h = 1; h < kModqBins; h++)
            out->modq_gap[r * VOSS_MODQ_MAX_GAP + 2 * h] = J.mq2[r * kModqBins + h];
    out->large_count = J.large_count;
    if (!out->large || !out->large_capacity) return;
    const size_t n = std::min<size_t>(J.large.size(), out->large_capacity);
    if (J.large.size() > (1u << 16) && n < J.large.size() / 4) {
        top_large(J.large, n, out->large);  // a few of many: select, not sort
        return;
    }
    sort_large(J.large);
    std::memcpy(out->large, J.large.data(), n * sizeof(voss_large_gap));
}

// Wheel-prime contribution to the residue counters (slot layout of the
// reference's modq_d, src/voss_v7_live.cu:195-232).
void wheel_modq(const JobHost& J, uint64_t* modq) {
    for (int o = 0; o < 8; o++) {
        modq[kW30[o] % 6] += J.mod30[o];
        modq[6 + kW30[o]] += J.mod30[o];
    }
}

// Windows of the grid w0 = k*W covering numbers [a, b] (a >= 7): k from
// ka = (a/30 - origin)/W to kb = (b/30 - origin)/W, each window non-empty
// (30*w0 <= b and 30*(w0+W) > a for every such k).
size_t plan_count(uint64_t a, uint64_t b, uint32_t W, uint64_t grid_origin) {
    if (a < 7) a = 7;
    if (a > b) return 0;
    return (size_t)((b / 30 - grid_origin) / W - (a / 30 - grid_origin) / W + 1);
}

// Writes the plan_count(a, b, W, origin) windows of [a, b] to out.  The
// base-prime count of consecutive windows only grows, so it is advanced
// with one pointer instead of a search per window.
void plan_fill(const voss_ctx* c, SegDesc* out, uint64_t a, uint64_t b, uint32_t job,
               uint32_t lseg, uint32_t W, uint64_t grid_origin) {
    if (a < 7) a = 7;
    if (a > b) return;
    const uint64_t ka = (a / 30 - grid_origin) / W, kb = (b / 30 - grid_origin) / W;
    size_t np = 0;
    const size_t nall = c->primes.size();
    const uint32_t ibig_w =
        (uint32_t)(std::lower_bound(c->primes.begin(), c->primes.end(), W) - c->primes.begin());
    for (uint64_t k = ka; k <= kb; k++) {
        const uint64_t w0 = grid_origin + k * W;
        const uint64_t lo = std::max<uint64_t>(a, 30 * w0);
        const uint64_t hi = std::min<uint64_t>(b, 30 * (w0 + W) - 1);
        uint64_t need = hi / 30 - w0 + 1;
        need = (need + 4095) / 4096 * 4096;
        SegDesc& d = *out++;
        d.w0 = w0;
        d.n_lo = lo;
        d.n_hi = hi;
        d.len = (uint32_t)std::min<uint64_t>(need, W);
        const uint64_t top = 30 * (w0 + d.len) - 1;
        if (np == 0) np = primes_upto(c, top);
        while (np < nall && (uint64_t)c->primes[np] * c->primes[np] <= top) np++;
        d.nprimes = (uint32_t)np;
        d.job = job;
        d.lseg = lseg;
        d.bucket = kNoBucket;
        d.ibig = d.len == W ? ibig_w
                            : (uint32_t)(std::lower_bound(c->primes.begin(), c->primes.end(),
                                                          d.len) - c->primes.begin());
    }
}

// Appends the windows of [a, b] to P.
void plan_range(const voss_ctx* c, Plan& P, uint64_t a, uint64_t b, uint32_t job,
                uint32_t lseg, uint32_t W, uint64_t grid_origin) {
    const size_t n = plan_count(a, b, W, grid_origin), at = P.desc.size();
    if (n == 0) return;
    P.desc.resize(at + n);
    plan_fill(c, P.desc.data() + at, a, b, job, lseg, W, grid_origin);
}

SieveArgs make_args(voss_ctx* c, ExecBufs& B, uint32_t ndesc, unsigned long long large_cap) {
    SieveArgs a{};
    a.desc = B.d_desc.p;
    a.ndesc = ndesc;
    a.tier_a_begin = kPresievePrimes;
    a.tier_a_end = (uint32_t)(std::lower_bound(c->primes.begin(), c->primes.end(),
                                               c->tune.tier_a_limit) - c->primes.begin());
    if (a.tier_a_end < a.tier_a_begin) a.tier_a_end = a.tier_a_begin;
    for (int h = 0; h < 3; h++)  // ceil(W / 4), ceil(W / 3), ceil(W / 2)
        a.ib_frac[h] = (uint32_t)(std::lower_bound(c->primes.begin(), c->primes.end(),
                                                   (c->tune.window_bytes + 3 - h) / (4 - h)) -
                                  c->primes.begin());
    a.restore0 = c->restore0;
    a.primes = c->d_primes.p;
    a.p32 = c->d_p32.p;
    const int block = sieve_block_threads((int)c->tune.window_bytes);
    a.pat = c->d_pat.p;
    for (int g = 0; g < kPresieveGroups; g++) {
        a.pat_off[g] = c->pat_off[g];
        a.pat_len[g] = c->pat_len[g];
        a.pat_step[g] = (2u * (uint32_t)block) % c->pat_len[g];  // a thread's next word pair
    }
    a.lg_thresh = c->d_lg_thresh.p;
    a.bucket_off = B.d_boff.p;
    a.bucket_cnt = B.d_bcur.p;
    a.bucket_ent = B.d_bent.p;
    a.emit_out = c->d_emit.p;
    a.emit_off = c->d_emit_off.p;
    a.rec = B.d_rec.p;
    a.acc = B.d_acc.p;
    a.large = B.d_large.p;
    a.large_cap = large_cap;
    a.large_n = B.d_flags.p;
    a.win_next = reinterpret_cast<uint32_t*>(B.d_flags.p + 3);
    return a;
}

// Room for the entries of one bucket (a window of W bytes) fed by primes
// with sum(1/p) = s.  A prime p has on average 8W/p multiples coprime to 30
// among the window's 30W numbers; the window's count is a sum of nearly
// independent per-prime counts whose variance is at most their mean, so
// mean + 8 sigma + 64 is essentially never exceeded.  If it is, the fill
// kernel reports it and the plan is re-run with exact counts.
uint32_t bucket_room(double s, uint32_t W, uint32_t pct) {
    const double e = 8.0 * W * s;
    const double r = (e + 8.0 * std::sqrt(e) + 64.0) * pct / 100.0;
    return (uint32_t)(((uint64_t)std::ceil(r) + 15) / 16 * 16);
}

// Estimated VALU work of a bucket task: per-prime setup (a class-mode
// prime also starts eight classes), class-mode hits, and walk steps (30
// span/p per walked prime).
double bucket_task_cost(const voss_ctx* c, const BucketTask& T, uint32_t W) {
    const double span = (double)T.nwin * W;
    return 25.0 * (4.0 * (double)(T.iw - T.i0) + (double)(T.i1 - T.iw)) +
           40.0 * span * (c->inv_prefix[T.iw] - c->inv_prefix[T.i0]) +
           210.0 * span * (c->inv_prefix[T.i1] - c->inv_prefix[T.iw]);
}

void make_rounds(const voss_ctx* c, const Plan& P, std::vector<Round>& rounds, double min_share) {
    const uint32_t W = c->tune.window_bytes;
    const uint32_t bl = c->tune.bucket_limit;
    const size_t ibl = (size_t)(std::lower_bound(c->primes.begin(), c->primes.end(), bl) -
                                c->primes.begin());
    std::vector<uint32_t> job_nwin(P.njobs, 0);
    for (const SegDesc& d : P.desc) job_nwin[d.job]++;
    auto bucketed = [&](const SegDesc& d) {
        return bl != 0xFFFFFFFFu && job_nwin[d.job] >= c->tune.bucket_min_windows &&
               d.nprimes > ibl;
    };
    // pass 1: round boundaries (the bucket budget is a running sum)
    struct Span {
        size_t off, end;
        bool any_bucket;
    };
    std::vector<Span> spans;
    size_t off = 0;
    uint32_t memo_np = ~0u;
    uint64_t memo_room = 0;
    while (off < P.desc.size()) {
        uint64_t est = 0;
        size_t end = off;
        bool any_bucket = false;
        while (end < P.desc.size() && end - off < kMaxDescPerLaunch) {
            const SegDesc& d = P.desc[end];
            if (bucketed(d)) {
                any_bucket = true;
                if (d.nprimes != memo_np) {  // consecutive windows mostly share it
                    memo_np = d.nprimes;
                    memo_room = bucket_room(c->inv_prefix[d.nprimes] - c->inv_prefix[ibl], W,
                                            c->tune.bucket_room_pct);
                }
                const uint64_t e = memo_room;
                if (end > off && est + e > kBucketEntryBudget * 15 / 16) break;
                est += e;
            }
            end++;
        }
        spans.push_back({off, end, any_bucket});
        off = end;
    }
    // pass 2: each round's buckets and tasks, rounds on several host threads
    rounds.assign(spans.size(), Round());
    parallel_for(spans.size(), [&](size_t r0, size_t r1) {
        for (size_t ri = r0; ri < r1; ri++) {
            Round& R = rounds[ri];
            R.off = spans[ri].off;
            const size_t end = spans[ri].end;
            R.nd = end - R.off;
            if (!spans[ri].any_bucket) continue;  // the plan's windows as they are (R.desc empty)
            const size_t off = R.off;
            R.desc.assign(P.desc.begin() + off, P.desc.begin() + end);
            std::vector<std::pair<double, BucketTask>> work;  // tasks with their estimated cost
            // buckets: consecutive windows of one job with the same w0 share one
            size_t i = 0;
            while (i < R.nd) {
                SegDesc& d = R.desc[i];
                if (!bucketed(d)) { i++; continue; }
                // a run of this job's windows (on its grid, step W)
                size_t j = i;
                uint32_t nb0 = R.nbuckets;
                uint64_t prev_w0 = ~0ull;
                uint32_t np_max = 0;
                std::vector<uint32_t> np_of;  // per bucket of the run: primes its windows need
                while (j < R.nd && R.desc[j].job == d.job && bucketed(R.desc[j])) {
                    SegDesc& e = R.desc[j];
                    if (e.w0 != prev_w0) {
                        if (prev_w0 != ~0ull && e.w0 != prev_w0 + W) break;
                        R.nbuckets++;
                        prev_w0 = e.w0;
                        np_of.push_back(0);
                    }
                    e.bucket = R.nbuckets - 1;
                    np_max = std::max(np_max, e.nprimes);
                    np_of.back() = std::max(np_of.back(), e.nprimes);
                    e.nprimes = (uint32_t)ibl;
                    j++;
                }
                const uint32_t nwin = R.nbuckets - nb0;
                const uint32_t room = bucket_room(c->inv_prefix[np_max] - c->inv_prefix[ibl], W,
                                                 c->tune.bucket_room_pct);
                R.room.insert(R.room.end(), nwin, room);
                R.room_total += (uint64_t)room * nwin;
                for (uint32_t t0 = 0; t0 < nwin; t0 += kBucketTileMax) {
                    const uint32_t tn = std::min<uint32_t>(kBucketTileMax, nwin - t0);
                    const uint32_t np_t = *std::max_element(np_of.begin() + t0, np_of.begin() + t0 + tn);
                    // primes beyond 1.5x the tile length: walk mode
                    const uint64_t walk_from = std::min<uint64_t>(3ull * tn * W / 2, 0xFFFFFFFFu);
                    const size_t iw = std::min<size_t>(
                        np_t, std::max<size_t>(ibl, (size_t)(std::lower_bound(
                                                                 c->primes.begin(), c->primes.end(),
                                                                 (uint32_t)walk_from) -
                                                             c->primes.begin())));
                    BucketTask T;
                    T.w0 = d.w0 + (uint64_t)t0 * W;
                    T.b0 = nb0 + t0;
                    T.nwin = tn;
                    T.i0 = (uint32_t)ibl;
                    T.iw = (uint32_t)iw;
                    T.i1 = np_t;
                    T.shared = 0;
                    work.push_back({bucket_task_cost(c, T, W), T});
                }
                i = j;
            }
            // a tile with more work than a workgroup's fair share is split into
            // prime chunks of about a quarter of that share (shared tasks)
            double total = 0;
            for (const auto& w : work) total += w.first;
            const double share = std::max(min_share, total / (2.0 * std::max(c->num_cus, 1)));  // two workgroups per CU
            std::vector<std::pair<double, BucketTask>> split;
            for (const auto& w : work) {
                if (w.first <= share) {
                    split.push_back(w);
                    continue;
                }
                const BucketTask& T = w.second;
                const uint32_t k = (uint32_t)std::min(256.0, std::ceil(4.0 * w.first / share));
                for (int walk = 0; walk < 2; walk++) {
                    const uint32_t a0 = walk ? T.iw : T.i0, a1 = walk ? T.i1 : T.iw;
                    if (a1 <= a0) continue;
                    const uint32_t step = std::max<uint32_t>(1, (a1 - a0 + k - 1) / k);
                    for (uint32_t x = a0; x < a1; x += step) {
                        BucketTask S = T;
                        S.i0 = x;
                        S.iw = walk ? x : std::min(a1, x + step);
                        S.i1 = std::min(a1, x + step);
                        S.shared = 1;
                        split.push_back({bucket_task_cost(c, S, W), S});
                    }
                }
            }
            work.swap(split);
            // largest tasks first, so the last ones to start are the short ones
            std::stable_sort(work.begin(), work.end(),
                             [](const std::pair<double, BucketTask>& x,
                                const std::pair<double, BucketTask>& y) { return x.first > y.first; });
            for (const auto& w : work) R.tasks.push_back(w.second);
        }
    }, 1);
}

int log2_window(const voss_ctx* c) {
    int l = 0;
    while ((1u << l) < c->tune.window_bytes) l++;
    return l;
}

// One-shot execution with the context's buffers.
int execute(voss_ctx* c, Plan&& P, std::vector<JobHost>& jobs, ExecTimes* times,
            std::vector<uint64_t>* counts, bool use_graph) {
    jobs.assign(P.njobs, JobHost());
    if (P.njobs == 0 || P.desc.empty()) {
        if (times) *times = ExecTimes();
        return VOSS_OK;
    }
    Executor X(c, &c->eb);
    X.P = std::move(P);
    X.plan_rounds();
    int rc = X.run(use_graph && c->graph_mode, times);
    if (!rc) rc = X.collect(jobs, counts);
    c->desc_pool.swap(X.P.desc);  // memory for the next plan
    return rc;
}

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

int check_result(voss_result* out) {
    if (!out) return fail(VOSS_EINVAL, "result pointer is NULL");
    if (out->large_capacity && !out->large) return fail(VOSS_EINVAL, "large buffer is NULL");
    return VOSS_OK;
}

// Zeroes every field but the caller's large-gap buffer.  The two big
// histograms are left to write_result, which fills them completely.
void clear_result(voss_result* out, bool histograms) {
    voss_large_gap* lg = out->large;
    const uint64_t cap = out->large_capacity;
    if (histograms) {
        std::memset(out, 0, sizeof(*out));
    } else {
        std::memset(out, 0, offsetof(voss_result, gaps));
        std::memset(reinterpret_cast<char*>(out) + offsetof(voss_result, large_count), 0,
                    sizeof(*out) - offsetof(voss_result, large_count));
    }
    out->large = lg;
    out->large_capacity = cap;
}

IntervalHead finish_interval_common(JobHost& J, uint64_t lo, uint64_t hi) {
    IntervalHead H;
    static const uint64_t small[3] = {2, 3, 5};
    for (uint64_t p : small) {
        if (p < lo || p >= hi) continue;
        if ((p & 3) == 1) H.c1++;
        if ((p & 3) == 3) H.c3++;
        H.modq[p % 6]++;
        H.modq[6 + p % 30]++;
        if (H.total) host_pair(J, H.last, p, true);
        else H.first = p;
        H.last = p;
        H.total++;
    }
    if (J.count) {
        if (H.total) host_pair(J, H.last, J.first, true);
        else H.first = J.first;
        H.last = J.last;
    }
    H.total += J.count;
    wheel_modq(J, H.modq);
    H.c1 += J.c1;
    H.c3 += J.count - J.c1;
    return H;
}

void finish_interval(JobHost& J, uint64_t lo, uint64_t hi, voss_result* out) {
    clear_result(out, false);  // write_result fills the histograms
    const IntervalHead H = finish_interval_common(J, lo, hi);
    out->total_primes = H.total;
    out->first_prime = H.first;
    out->last_prime = H.last;
    out->c1 = H.c1;
    out->c3 = H.c3;
    std::memcpy(out->modq, H.modq, sizeof(H.modq));
    write_result(J, out);
}

void finish_interval(JobHost& J, uint64_t lo, uint64_t hi, voss_interval_result* out) {
    voss_large_gap* lg = out->large;
    const uint64_t cap = out->large_capacity;
    std::memset(out, 0, sizeof(*out));
    out->large = lg;
    out->large_capacity = cap;
    out->lo = lo;
    out->hi = hi;
    const IntervalHead H = finish_interval_common(J, lo, hi);
    out->total_primes = H.total;
    out->first_prime = H.first;
    out->last_prime = H.last;
    out->c1 = H.c1;
    out->c3 = H.c3;
    std::memcpy(out->modq, H.modq, sizeof(H.modq));
    J.init();
    out->gaps[0] = (uint64_t)J.gap1;
    for (int h = 1; h < kHistBins; h++) out->gaps[h] = (uint64_t)J.hist2[h];
    for (int i = 0; i < kModqRows * kModqBins; i++) out->modq_gap[i] = (uint64_t)J.mq2[i];
    out->large_count = J.large_count;
    sort_large(J.large);
    if (out->large && out->large_capacity) {
        const size_t n = std::min<size_t>(J.large.size(), out->large_capacity);
        std::memcpy(out->large, J.large.data(), n * sizeof(voss_large_gap));
    }
}

}  // namespace core
}  // namespace voss

extern "C" {

const char* voss_last_error(void) { return g_err.c_str(); }

int voss_create(int device, voss_ctx** out) {
    Admit admit_(nullptr, true);
    if (!out) return fail(VOSS_EINVAL, "out is NULL");
    *out = nullptr;
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(VOSS_EINVAL, "no such device");
    voss_ctx* c = new voss_ctx();
    c->device = device;
    int rc = set_device(c);
    hipDeviceProp_t prop;
    if (rc == VOSS_OK) {
        hipError_t e = hipGetDeviceProperties(&prop, device);
        if (e != hipSuccess) rc = hip_fail(e, "hipGetDeviceProperties");
    }
    if (rc == VOSS_OK) {
        c->name = prop.name[0] ? prop.name : prop.gcnArchName;
        c->num_cus = prop.multiProcessorCount;
        c->total_gb = prop.totalGlobalMem / 1e9;
        hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreate(&c->ev0);
        if (e == hipSuccess) e = hipEventCreate(&c->ev1);
        if (e == hipSuccess) e = hipEventCreate(&c->ev2);
        if (e != hipSuccess) rc = hip_fail(e, "context setup");
        if (rc == VOSS_OK && (e = mr_scratch_address(device, &c->mr_scratch)) != hipSuccess)
            rc = hip_fail(e, "hipGetSymbolAddress(scratch)");
    }
    if (rc == VOSS_OK) rc = build_patterns(c);
    if (rc == VOSS_OK) rc = build_lg_thresh(c);
    if (rc == VOSS_OK) rc = ensure_primes(c, 1ull << 32);
    if (rc != VOSS_OK) {
        voss_destroy(c);
        return rc;
    }
    *out = c;
    return VOSS_OK;
}

void voss_destroy(voss_ctx* c) {
    if (!c) return;
    Admit admit_(nullptr, true);
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    c->d_primes.release();
    c->d_p32.release();
    c->d_pat.release();
    c->d_lg_thresh.release();
    c->eb.release();
    for (hipEvent_t e : c->evs) (void)hipEventDestroy(e);
    c->evs.clear();
    c->d_emit.release();
    c->d_emit_off.release();
    c->d_mr_n.release();
    c->d_mr_v.release();
    c->d_surv.release();
    c->d_nsurv.release();
    c->d_vf_pa.release();
    c->d_vf_gap.release();
    c->d_vf_ok.release();
    if (c->serve_cache) c->free_serve_cache(c->serve_cache);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->ev2) (void)hipEventDestroy(c->ev2);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int voss_device_name(voss_ctx* c, char* buf, size_t cap, int* num_cus, double* total_gb) {
    if (!c) return fail(VOSS_EINVAL, "ctx is NULL");
    if (buf && cap) {
        std::snprintf(buf, cap, "%s", c->name.c_str());
    }
    if (num_cus) *num_cus = c->num_cus;
    if (total_gb) *total_gb = c->total_gb;
    return VOSS_OK;
}

int voss_set_tuning(voss_ctx* c, const voss_tuning* t) {
    VOSS_ADMIT(c);
    if (!c || !t) return fail(VOSS_EINVAL, "NULL argument");
    voss_tuning n = c->tune;
    if (t->window_bytes) {
        if (sieve_block_threads((int)t->window_bytes) == 0)
            return fail(VOSS_EINVAL, "window_bytes must be 32768, 65536 or 131072");
        n.window_bytes = t->window_bytes;
    }
    if (t->wg_per_cu) n.wg_per_cu = t->wg_per_cu;
    if (t->tier_a_limit) n.tier_a_limit = t->tier_a_limit;
    if (t->bucket_limit) {
        if (t->bucket_limit < n.window_bytes && t->bucket_limit != 0xFFFFFFFFu)
            return fail(VOSS_EINVAL, "bucket_limit must be at least the window length");
        n.bucket_limit = t->bucket_limit;
    }
    if (t->bucket_min_windows) n.bucket_min_windows = t->bucket_min_windows;
    if (t->bucket_room_pct) {
        if (t->bucket_room_pct > 1000) return fail(VOSS_EINVAL, "bucket_room_pct beyond 1000");
        n.bucket_room_pct = t->bucket_room_pct;
    }
    c->tune = n;
    return VOSS_OK;
}

int voss_autotune(voss_ctx* c, const uint64_t* lo, const uint64_t* hi, size_t n,
                  size_t max_intervals, int repeats, voss_tuning* best, double* best_seconds) {
    VOSS_ADMIT(c);
    if (!c || (n && (!lo || !hi))) return fail(VOSS_EINVAL, "NULL argument");
    if (n == 0) return fail(VOSS_EINVAL, "no intervals to tune on");
    if (repeats < 1) repeats = 1;
    if (max_intervals == 0) max_intervals = n;
    // evenly spread sample of the workload
    std::vector<uint64_t> slo, shi;
    const size_t m = std::min(n, max_intervals);
    for (size_t k = 0; k < m; k++) {
        const size_t i = k * n / m;
        slo.push_back(lo[i]);
        shi.push_back(hi[i]);
    }
    std::vector<voss_interval_result> res(m);
    auto time_it = [&](const voss_tuning& t, double* sec) -> int {
        int rc = voss_set_tuning(c, &t);
        if (rc) return rc;
        double bestt = 1e300;
        for (int r = 0; r <= repeats; r++) {
            for (auto& x : res) { x.large = nullptr; x.large_capacity = 0; }
            const double t0 = now_s();
            if ((rc = voss_run_batch(c, slo.data(), shi.data(), m, res.data(), nullptr))) return rc;
            const double dt = now_s() - t0;
            if (r > 0) bestt = std::min(bestt, dt);  // the first run warms up
        }
        *sec = bestt;
        return VOSS_OK;
    };
    voss_tuning cur = c->tune;
    double cur_s = 0;
    int rc = time_it(cur, &cur_s);
    if (rc) return rc;
    // (window bytes, workgroups per CU): one CU holds one 128 KiB window's
    // workgroup, two 64 KiB ones or three 32 KiB ones (their LDS: 162,880,
    // 80,688 and 45,872 bytes of the CU's 163,840; the occupancy the runtime
    // reports for the three sieve kernels)
    const std::vector<std::pair<uint32_t, uint32_t>> shapes = {
        {131072, 1}, {65536, 2}, {65536, 1}, {32768, 3}};
    const std::vector<uint32_t> tiers = {512, 768, 1024, 2048};
    const std::vector<uint32_t> blimits = {1u << 17, 1u << 18, 1u << 19, 1u << 20, 1u << 21,
                                           1u << 22, 1u << 23, 1u << 25, 0xFFFFFFFFu};
    const std::vector<uint32_t> bmins = {1, 2, 4, 8};
    for (int pass = 0; pass < 2; pass++) {
        bool moved = false;
        for (int dim = 0; dim < 4; dim++) {
            const size_t nv = dim == 0 ? shapes.size() : dim == 1 ? tiers.size()
                            : dim == 2 ? blimits.size() : bmins.size();
            for (size_t k = 0; k < nv; k++) {
                voss_tuning t = cur;
                if (dim == 0) {
                    t.window_bytes = shapes[k].first;
                    t.wg_per_cu = shapes[k].second;
                    if (t.bucket_limit < t.window_bytes) t.bucket_limit = t.window_bytes;
                } else if (dim == 1) {
                    t.tier_a_limit = tiers[k];
                } else if (dim == 2) {
                    if (blimits[k] < t.window_bytes) continue;
                    t.bucket_limit = blimits[k];
                } else {
                    t.bucket_min_windows = bmins[k];
                }
                if (std::memcmp(&t, &cur, sizeof t) == 0) continue;
                double sec = 0;
                if ((rc = time_it(t, &sec))) return rc;
                if (sec < cur_s * 0.98) {  // ignore changes within the noise
                    cur = t;
                    cur_s = sec;
                    moved = true;
                }
            }
        }
        if (!moved) break;
    }
    // the bucket limit trades per-window prime setup against bucket
    // entries, a balance that moves with the batch's size: when the sample
    // was a part of the batch, the limit and its two neighbours are timed
    // once more on all n intervals
    if (m < n && cur.bucket_limit != 0xFFFFFFFFu) {
        std::vector<voss_interval_result> full(n);
        auto time_full = [&](const voss_tuning& t, double* sec) -> int {
            int rc2 = voss_set_tuning(c, &t);
            if (rc2) return rc2;
            double bestt = 1e300;
            for (int r = 0; r <= repeats; r++) {
                for (auto& x : full) { x.large = nullptr; x.large_capacity = 0; }
                const double t0 = now_s();
                if ((rc2 = voss_run_batch(c, lo, hi, n, full.data(), nullptr))) return rc2;
                if (r > 0) bestt = std::min(bestt, now_s() - t0);
            }
            *sec = bestt;
            return VOSS_OK;
        };
        double full_s = 0;
        if ((rc = time_full(cur, &full_s))) return rc;
        const voss_tuning base = cur;
        for (const uint32_t bl : {base.bucket_limit / 2, base.bucket_limit * 2}) {
            if (bl < base.window_bytes || bl > (1u << 30)) continue;
            voss_tuning t = base;
            t.bucket_limit = bl;
            double sec = 0;
            if ((rc = time_full(t, &sec))) return rc;
            if (sec < full_s * 0.98) {
                cur = t;
                full_s = sec;
            }
        }
        cur_s = full_s;
    }
    if ((rc = voss_set_tuning(c, &cur))) return rc;
    if (best) *best = cur;
    if (best_seconds) *best_seconds = cur_s;
    return VOSS_OK;
}

int voss_get_tuning(voss_ctx* c, voss_tuning* t) {
    if (!c || !t) return fail(VOSS_EINVAL, "NULL argument");
    *t = c->tune;
    return VOSS_OK;
}

}  // extern "C"

namespace {

int plain_args(voss_ctx* c, uint64_t N, uint64_t seg_num, voss_result* out) {
    if (!c) return fail(VOSS_EINVAL, "ctx is NULL");
    int rc = check_result(out);
    if (rc) return rc;
    if (N < 5) return fail(VOSS_EINVAL, "N must be >= 5");
    if (seg_num == 0 || seg_num % 30 != 0)
        return fail(VOSS_EINVAL, "SEG_NUM must be a positive multiple of 30");
    if (N > (1ull << 62)) return fail(VOSS_EINVAL, "N beyond 2^62");
    return set_device(c);
}

Plan plain_plan(voss_ctx* c, uint64_t N, uint64_t seg_num, uint64_t s0, uint64_t s1) {
    Plan P;
    P.desc.swap(c->desc_pool);  // reuse the last plan's memory
    P.desc.clear();
    P.njobs = 1;
    const uint32_t W = c->tune.window_bytes;
    // window counts per logical segment, then the segments planned in
    // parallel into their places (a plain run to 1e14: 3,334 segments,
    // 2.5e7 windows)
    const size_t nseg = (size_t)(s1 - s0);
    std::vector<size_t> at(nseg + 1, 0);
    auto range_of = [&](size_t i, uint64_t& a, uint64_t& b) {
        const uint64_t s = s0 + i;
        a = s * seg_num + 1;
        b = std::min<uint64_t>((s + 1) * seg_num, N);
    };
    for (size_t i = 0; i < nseg; i++) {
        uint64_t a, b;
        range_of(i, a, b);
        at[i + 1] = at[i] + plan_count(a, b, W, 0);
    }
    P.desc.resize(at[nseg]);
    // pieces of at most kPiece windows (a segment's windows k = ka..kb of
    // the grid 30*W*k; a piece clips [a, b] to its windows, which plan_fill
    // then produces exactly as for the whole segment), filled in parallel
    constexpr size_t kPiece = 2048;
    struct Piece {
        uint64_t a, b;
        size_t at;
        uint32_t seg;
    };
    std::vector<Piece> pieces;
    pieces.reserve(at[nseg] / kPiece + nseg);
    for (size_t i = 0; i < nseg; i++) {
        uint64_t a, b;
        range_of(i, a, b);
        for (size_t k = 0; at[i] + k < at[i + 1]; k += kPiece) {
            uint64_t pa = a, pb = b;
            if (k) {  // the piece starts at its first window
                const uint64_t ka = (std::max<uint64_t>(a, 7) / 30) / W;
                pa = std::max<uint64_t>(a, 30 * (ka + k) * W);
            }
            if (at[i] + k + kPiece < at[i + 1]) {  // and ends at its last
                const uint64_t ka = (std::max<uint64_t>(a, 7) / 30) / W;
                pb = std::min<uint64_t>(b, 30 * (ka + k + kPiece) * W - 1);
            }
            pieces.push_back({pa, pb, at[i] + k, (uint32_t)(s0 + i)});
        }
    }
    parallel_for(pieces.size(), [&](size_t i0, size_t i1) {
        for (size_t i = i0; i < i1; i++)
            plan_fill(c, P.desc.data() + pieces[i].at, pieces[i].a, pieces[i].b, 0, pieces[i].seg,
                      W, 0);
    }, 1);
    return P;
}

// Logical segments [s0, s1) of the plain run: primes >= 7 only, no head
// pairs.  Leaves J with first/last prime of the range.
int plain_range(voss_ctx* c, uint64_t N, uint64_t seg_num, uint64_t s0, uint64_t s1,
                voss_result* out, JobHost& J, ExecTimes& tm, uint64_t* nwin) {
    const uint32_t W = c->tune.window_bytes;
    int rc = ensure_primes(c, (N / 30 / W + 1) * W * 30);
    if (rc) return rc;
    Plan P = plain_plan(c, N, seg_num, s0, s1);
    *nwin = P.desc.size();
    std::vector<JobHost> jobs;
    std::vector<uint64_t> counts;
    (void)out;
    if ((rc = execute(c, std::move(P), jobs, &tm, &counts))) return rc;
    if (s0 == 0 && s1 == (N + seg_num - 1) / seg_num) {
        c->last_N = N;
        c->last_seg = seg_num;
        c->last_W = W;
        c->last_counts.swap(counts);
    }
    J = std::move(jobs[0]);
    return VOSS_OK;
}

// Checks a batch and sizes the base primes for it.
int batch_args(voss_ctx* c, const uint64_t* lo, const uint64_t* hi, size_t n) {
    if (!c) return fail(VOSS_EINVAL, "ctx is NULL");
    if (n && (!lo || !hi)) return fail(VOSS_EINVAL, "NULL argument");
    if (n > 0xFFFFFFFFull) return fail(VOSS_EINVAL, "too many intervals");
    uint64_t maxhi = 0;
    for (size_t i = 0; i < n; i++) {
        if (hi[i] < lo[i]) return fail(VOSS_EINVAL, "hi < lo");
        if (hi[i] > (1ull << 62)) return fail(VOSS_EINVAL, "hi beyond 2^62");
        maxhi = std::max(maxhi, hi[i]);
    }
    int rc = set_device(c);
    if (rc) return rc;
    return ensure_primes(c, maxhi / 30 * 30 + 30ull * c->tune.window_bytes + 30);
}

int check_interval_out(const voss_interval_result* out, size_t n) {
    if (n && !out) return fail(VOSS_EINVAL, "NULL argument");
    for (size_t i = 0; i < n; i++)
        if (out[i].large_capacity && !out[i].large) return fail(VOSS_EINVAL, "large buffer is NULL");
    return VOSS_OK;
}

// Job i = interval i, each on its own window grid.
Plan plan_batch(voss_ctx* c, const uint64_t* lo, const uint64_t* hi, size_t n) {
    Plan P;
    P.desc.swap(c->desc_pool);
    P.desc.clear();
    P.njobs = (uint32_t)n;
    const uint32_t W = c->tune.window_bytes;
    auto plan_jobs = [&](Plan& Q, size_t i0, size_t i1) {
        for (size_t i = i0; i < i1; i++) {
            if (hi[i] <= 7) continue;
            const uint64_t a = std::max<uint64_t>(lo[i], 7), b = hi[i] - 1;
            plan_range(c, Q, a, b, (uint32_t)i, 0, W, (a / 30) & ~4095ull);
        }
    };
    if (n < 2048) {
        plan_jobs(P, 0, n);
        return P;
    }
    // large batches: job ranges planned on several host threads, then
    // concatenated in job order
    std::mutex m;
    std::vector<std::pair<size_t, Plan>> parts;
    parallel_for(n, [&](size_t i0, size_t i1) {
        Plan Q;
        plan_jobs(Q, i0, i1);
        std::lock_guard<std::mutex> g(m);
        parts.emplace_back(i0, std::move(Q));
    }, 256);
    std::sort(parts.begin(), parts.end(),
              [](const std::pair<size_t, Plan>& x, const std::pair<size_t, Plan>& y) {
                  return x.first < y.first;
              });
    size_t total = 0;
    for (const auto& q : parts) total += q.second.desc.size();
    P.desc.resize(total);
    size_t at = 0;
    for (const auto& q : parts) {
        std::memcpy(P.desc.data() + at, q.second.desc.data(), q.second.desc.size() * sizeof(SegDesc));
        at += q.second.desc.size();
    }
    return P;
}

void batch_stats(voss_result* stats, const ExecTimes& tm, double t0, uint64_t nwin) {
    if (!stats) return;
    clear_result(stats);
    stats->seconds = now_s() - t0;
    stats->device_ms = tm.device_ms;
    stats->sieve_ms = tm.sieve_ms;
    stats->windows = nwin;
    stats->bucket_ms = tm.bucket_ms;
    stats->bucket_entries = tm.bucket_entries;
}

int finish_batch(std::vector<JobHost>& jobs, const uint64_t* lo, const uint64_t* hi, size_t n,
                 voss_interval_result* out) {
    parallel_for(n, [&](size_t i0, size_t i1) {
        for (size_t i = i0; i < i1; i++) finish_interval(jobs[i], lo[i], hi[i], &out[i]);
    }, 256);
    for (size_t i = 0; i < n; i++)
        if (jobs[i].overflow) return fail(VOSS_ERANGE, "prime gap >= 2048");
    return VOSS_OK;
}

int run_batch(voss_ctx* c, const uint64_t* lo, const uint64_t* hi, size_t n,
              voss_interval_result* out, voss_result* stats, bool use_graph) {
    int rc = batch_args(c, lo, hi, n);
    if (rc || (rc = check_interval_out(out, n))) return rc;
    const double t0 = now_s();
    Plan P = plan_batch(c, lo, hi, n);
    const uint64_t nwin = P.desc.size();
    std::vector<JobHost> jobs;
    ExecTimes tm;
    if ((rc = execute(c, std::move(P), jobs, &tm, nullptr, use_graph))) return rc;
    if ((rc = finish_batch(jobs, lo, hi, n, out))) return rc;
    batch_stats(stats, tm, t0, nwin);
    return VOSS_OK;
}

void stamp_times(voss_result* out, const ExecTimes& tm, double t0, uint64_t nwin) {
    out->seconds = now_s() - t0;
    out->device_ms = tm.device_ms;
    out->sieve_ms = tm.sieve_ms;
    out->windows = nwin;
    out->bucket_ms = tm.bucket_ms;
    out->bucket_entries = tm.bucket_entries;
}

}  // namespace

extern "C" {

int voss_run_plain(voss_ctx* c, uint64_t N, uint64_t seg_num, voss_result* out) {
    VOSS_ADMIT(c);
    int rc = plain_args(c, N, seg_num, out);
    if (rc) return rc;
    const double t0 = now_s();
    JobHost J;
    ExecTimes tm;
    uint64_t nwin = 0;
    const uint64_t num_seg = (N + seg_num - 1) / seg_num;  // :342
    if ((rc = plain_range(c, N, seg_num, 0, num_seg, out, J, tm, &nwin))) return rc;
    clear_result(out, false);  // write_result fills the histograms
    // (5, first wheel prime) is the reference's host boundary gap of
    // segment 0 (:559-562): main histogram only
    if (J.count) host_pair(J, 5, J.first, false);
    J.init();
    J.gap1 += 1;      // (2,3)  :634
    J.hist2[1] += 1;  // (3,5)  :635
    if (J.overflow) return fail(VOSS_ERANGE, "prime gap >= 2048");
    out->total_primes = 3 + J.count;
    out->first_prime = 2;
    out->last_prime = J.count ? J.last : 5;
    out->c1 = J.c1;
    out->c3 = J.count - J.c1;
    wheel_modq(J, out->modq);
    write_result(J, out);
    stamp_times(out, tm, t0, nwin);
    return VOSS_OK;
}

int voss_run_plain_range(voss_ctx* c, uint64_t N, uint64_t seg_num, uint64_t seg_begin,
                         uint64_t seg_end, voss_result* out) {
    VOSS_ADMIT(c);
    int rc = plain_args(c, N, seg_num, out);
    if (rc) return rc;
    const uint64_t num_seg = (N + seg_num - 1) / seg_num;
    if (seg_begin > seg_end || seg_end > num_seg)
        return fail(VOSS_EINVAL, "segment range outside [0, NUM_SEG]");
    const double t0 = now_s();
    JobHost J;
    ExecTimes tm;
    uint64_t nwin = 0;
    if ((rc = plain_range(c, N, seg_num, seg_begin, seg_end, out, J, tm, &nwin))) return rc;
    if (J.overflow) return fail(VOSS_ERANGE, "prime gap >= 2048");
    clear_result(out, false);  // write_result fills the histograms
    out->total_primes = J.count;
    out->first_prime = J.count ? J.first : 0;
    out->last_prime = J.count ? J.last : 0;
    out->c1 = J.c1;
    out->c3 = J.count - J.c1;
    wheel_modq(J, out->modq);
    write_result(J, out);
    stamp_times(out, tm, t0, nwin);
    return VOSS_OK;
}

int voss_run_interval(voss_ctx* c, uint64_t lo, uint64_t hi, voss_result* out) {
    VOSS_ADMIT(c);
    if (!c) return fail(VOSS_EINVAL, "ctx is NULL");
    int rc = check_result(out);
    if (rc) return rc;
    if (hi < lo) return fail(VOSS_EINVAL, "hi < lo");
    if (hi > (1ull << 62)) return fail(VOSS_EINVAL, "hi beyond 2^62");
    if ((rc = set_device(c))) return rc;
    const uint32_t W = c->tune.window_bytes;
    if ((rc = ensure_primes(c, hi / 30 * 30 + 30ull * W + 30))) return rc;
    const double t0 = now_s();
    Plan P;
    P.njobs = 1;
    if (hi > 7) {
        const uint64_t a = std::max<uint64_t>(lo, 7), b = hi - 1;
        plan_range(c, P, a, b, 0, 0, W, (a / 30) & ~4095ull);
    }
    std::vector<JobHost> jobs;
    ExecTimes tm;
    const uint64_t nwin = P.desc.size();
    if ((rc = execute(c, std::move(P), jobs, &tm))) return rc;
    finish_interval(jobs[0], lo, hi, out);
    if (jobs[0].overflow) return fail(VOSS_ERANGE, "prime gap >= 2048");
    out->seconds = now_s() - t0;
    out->device_ms = tm.device_ms;
    out->sieve_ms = tm.sieve_ms;
    out->windows = nwin;
    out->bucket_ms = tm.bucket_ms;
    out->bucket_entries = tm.bucket_entries;
    return VOSS_OK;
}

int voss_run_batch(voss_ctx* c, const uint64_t* lo, const uint64_t* hi, size_t n,
                   voss_interval_result* out, voss_result* stats) {
    VOSS_ADMIT(c);
    return run_batch(c, lo, hi, n, out, stats, c ? c->graph_mode != 0 : false);
}

int voss_run_batch_plain(voss_ctx* c, const uint64_t* lo, const uint64_t* hi, size_t n,
                         voss_interval_result* out, voss_result* stats) {
    VOSS_ADMIT(c);
    if (!c) return fail(VOSS_EINVAL, "ctx is NULL");
    if (n && (!lo || !hi || !out)) return fail(VOSS_EINVAL, "NULL argument");
    const double t0 = now_s();
    double dev = 0, sv = 0;
    uint64_t win = 0;
    for (size_t i = 0; i < n; i++) {
        voss_result st;
        int rc = run_batch(c, lo + i, hi + i, 1, out + i, &st, false);
        if (rc) return rc;
        dev += st.device_ms;
        sv += st.sieve_ms;
        win += st.windows;
    }
    if (stats) {
        clear_result(stats);
        stats->seconds = now_s() - t0;
        stats->device_ms = dev;
        stats->sieve_ms = sv;
        stats->windows = win;
    }
    return VOSS_OK;
}

}  // extern "C"

struct voss_batch {
    voss_ctx* c = nullptr;
    voss_tuning tune{};
    std::vector<uint64_t> lo, hi;
    uint64_t nwin = 0;
    ExecBufs B;
    std::unique_ptr<Executor> X;
    std::vector<std::unique_ptr<DevBuf<uint32_t>>> mr_scratch;  // survivors + counter per MR batch
    ~voss_batch() {
        for (auto& m : mr_scratch) m->release();
    }
};

extern "C" {

int voss_batch_prepare(voss_ctx* c, const uint64_t* lo, const uint64_t* hi, size_t n,
                       voss_batch** out) {
    VOSS_ADMIT(c);
    if (!out) return fail(VOSS_EINVAL, "out is NULL");
    *out = nullptr;
    int rc = batch_args(c, lo, hi, n);
    if (rc) return rc;
    std::unique_ptr<voss_batch> b(new voss_batch());
    b->c = c;
    b->tune = c->tune;
    b->lo.assign(lo, lo + n);
    b->hi.assign(hi, hi + n);
    b->X.reset(new Executor(c, &b->B));
    b->X->P = plan_batch(c, lo, hi, n);
    b->nwin = b->X->P.desc.size();
    b->X->plan_rounds();
    *out = b.release();
    return VOSS_OK;
}

int voss_batch_run(voss_batch* b, voss_interval_result* out, voss_result* stats) {
    VOSS_ADMIT(b ? b->c : nullptr);
    if (!b) return fail(VOSS_EINVAL, "batch is NULL");
    voss_ctx* c = b->c;
    const size_t n = b->lo.size();
    int rc = check_interval_out(out, n);
    if (rc) return rc;
    if (std::memcmp(&c->tune, &b->tune, sizeof(voss_tuning)) != 0)
        return fail(VOSS_EINVAL, "tuning changed since voss_batch_prepare");
    if ((rc = set_device(c))) return rc;
    const double t0 = now_s();
    std::vector<JobHost> jobs;
    ExecTimes tm;
    if (b->X->P.desc.empty() && b->X->mr.empty()) {
        jobs.assign(n, JobHost());
    } else {
        if ((rc = b->X->run(true, &tm))) return rc;
        if ((rc = b->X->collect(jobs, nullptr))) return rc;
    }
    if ((rc = finish_batch(jobs, b->lo.data(), b->hi.data(), n, out))) return rc;
    batch_stats(stats, tm, t0, b->nwin);
    return VOSS_OK;
}

int voss_batch_add_mr(voss_batch* b, const uint64_t* d_n, size_t count, uint8_t* d_verdict) {
    VOSS_ADMIT(b ? b->c : nullptr);
    if (!b) return fail(VOSS_EINVAL, "batch is NULL");
    if (count && (!d_n || !d_verdict)) return fail(VOSS_EINVAL, "NULL buffer");
    if (count >= 0xFFFFFFFFull) return fail(VOSS_EINVAL, "count must be below 2^32");
    if (count == 0) return VOSS_OK;
    int rc = set_device(b->c);
    if (rc || (rc = b->X->ensure_side())) return rc;
    std::unique_ptr<DevBuf<uint32_t>> scratch(new DevBuf<uint32_t>());
    HIPCHK(scratch->reserve(2 * count + 3));  // two survivor lists, then their counters
    b->X->mr.push_back({d_n, count, d_verdict, scratch->p, scratch->p + 2 * count});
    b->mr_scratch.push_back(std::move(scratch));
    b->X->drop_graph();  // the next run captures the sequence with this batch in it
    return VOSS_OK;
}

void voss_batch_free(voss_batch* b) {
    if (!b) return;
    Admit admit_(nullptr, true);
    (void)hipSetDevice(b->c->device);
    (void)hipStreamSynchronize(b->c->stream);
    b->X.reset();
    b->B.release();
    delete b;
}

int voss_set_graph_mode(voss_ctx* c, int enable) {
    VOSS_ADMIT(c);
    if (!c) return fail(VOSS_EINVAL, "ctx is NULL");
    c->graph_mode = enable ? 1 : 0;
    return VOSS_OK;
}

int voss_sample_plain(voss_ctx* c, uint64_t N, uint64_t seg_num, int32_t* gap_seq,
                      uint64_t* pos_seq, uint64_t* rows_out) {
    VOSS_ADMIT(c);
    if (!gap_seq || !pos_seq) return fail(VOSS_EINVAL, "NULL buffer");
    int rc;
    const uint32_t W = c ? c->tune.window_bytes : 0;
    if (!c || c->last_N != N || c->last_seg != seg_num || c->last_W != W) {
        // per-window counts come from a plain run of the same range
        voss_result* tmp = (voss_result*)std::calloc(1, sizeof(voss_result));
        if (!tmp) return fail(VOSS_ENOMEM, "host allocation");
        rc = voss_run_plain(c, N, seg_num, tmp);
        std::free(tmp);
        if (rc) return rc;
    }
    if ((rc = set_device(c))) return rc;
    std::memset(gap_seq, 0, sizeof(int32_t) * VOSS_GAP_SEQ_SIZE);
    std::memset(pos_seq, 0, sizeof(uint64_t) * VOSS_GAP_SEQ_SIZE);
    const uint64_t num_seg = (N + seg_num - 1) / seg_num;
    const Plan P = plain_plan(c, N, seg_num, 0, num_seg);
    const std::vector<uint64_t>& cnt = c->last_counts;
    if (cnt.size() != P.desc.size()) return fail(VOSS_EINVAL, "stale window counts");
    const int64_t per_seg = VOSS_GAP_SEQ_SIZE / (int64_t)num_seg;  // :529
    struct Pick { uint64_t seg, skip, n; size_t emit0; };
    std::vector<Pick> picks;
    std::vector<SegDesc> emit_desc;
    std::vector<unsigned long long> emit_off;
    unsigned long long total = 0;
    size_t d = 0;
    for (uint64_t sg = 0; sg < num_seg; sg++) {
        const size_t d0 = d;
        uint64_t n_pos = 0;
        while (d < P.desc.size() && P.desc[d].lseg == (uint32_t)sg) n_pos += cnt[d++];
        if (per_seg <= 0 || n_pos <= 1) continue;
        const int64_t start = (int64_t)(n_pos / 2);
        const int64_t avail = (int64_t)n_pos - start - 1;
        const int64_t sample_n = avail < per_seg ? avail : per_seg;
        if (sample_n <= 1) continue;
        if ((int64_t)sg * per_seg + sample_n > VOSS_GAP_SEQ_SIZE) continue;
        // windows holding in-segment prime indices start .. start+sample_n
        uint64_t before = 0;
        size_t k = d0;
        while (before + cnt[k] <= (uint64_t)start) before += cnt[k++];
        Pick pk{sg, (uint64_t)start - before, (uint64_t)sample_n + 1, emit_off.size()};
        uint64_t covered = before;
        const size_t first_emit = emit_desc.size();
        while (k < d && covered < (uint64_t)start + (uint64_t)sample_n + 1) {
            emit_desc.push_back(P.desc[k]);
            emit_off.push_back(total);
            total += cnt[k];
            covered += cnt[k];
            k++;
        }
        pk.emit0 = first_emit;
        picks.push_back(pk);
    }
    uint64_t rows = 0;
    if (!emit_desc.empty()) {
        const uint32_t nd = (uint32_t)emit_desc.size();
        ExecBufs& B = c->eb;
        HIPCHK(B.d_desc.reserve(nd));
        HIPCHK(B.d_rec.reserve(nd));
        HIPCHK(c->d_emit_off.reserve(nd));
        HIPCHK(c->d_emit.reserve(std::max<unsigned long long>(total, 1)));
        HIPCHK(B.d_acc.reserve(1));
        HIPCHK(B.d_large.reserve(1));
        HIPCHK(B.d_flags.reserve(4));
        HIPCHK(hipMemcpyAsync(B.d_desc.p, emit_desc.data(), nd * sizeof(SegDesc),
                              hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(c->d_emit_off.p, emit_off.data(), nd * 8, hipMemcpyHostToDevice,
                              c->stream));
        HIPCHK(launch_init_acc(B.d_acc.p, 1, c->stream));
        SieveArgs a = make_args(c, B, nd, 1);
        HIPCHK(launch_fill_u32(a.win_next, 1, 0, c->stream));
        HIPCHK(launch_sieve_emit(a, (int)W, (uint32_t)c->num_cus * c->tune.wg_per_cu, c->stream));
        std::vector<unsigned long long> primes(total);
        HIPCHK(hipMemcpyAsync(primes.data(), c->d_emit.p, total * 8, hipMemcpyDeviceToHost,
                              c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        for (const Pick& pk : picks) {
            const unsigned long long* v = primes.data() + emit_off[pk.emit0] + pk.skip;
            const int64_t off = (int64_t)pk.seg * per_seg;
            for (uint64_t i = 0; i + 1 < pk.n; i++) {
                gap_seq[off + (int64_t)i] = (int32_t)(v[i + 1] - v[i]);  // :250
                pos_seq[off + (int64_t)i] = v[i];                        // :251
                rows++;
            }
        }
    }
    if (rows_out) *rows_out = rows;
    return VOSS_OK;
}

int voss_mr_batch_device(voss_ctx* c, const uint64_t* d_n, size_t count, uint8_t* d_verdict,
                         double* kernel_ms) {
    VOSS_ADMIT(c);
    if (!c) return fail(VOSS_EINVAL, "ctx is NULL");
    if (count && (!d_n || !d_verdict)) return fail(VOSS_EINVAL, "NULL buffer");
    if (count >= 0xFFFFFFFFull) return fail(VOSS_EINVAL, "count must be below 2^32");
    int rc;
    if ((rc = set_device(c))) return rc;
    if (count == 0) {
        if (kernel_ms) *kernel_ms = 0;
        return VOSS_OK;
    }
    HIPCHK(c->d_surv.reserve(2 * count));
    HIPCHK(c->d_nsurv.reserve(3));
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    HIPCHK(launch_mr(d_n, count, d_verdict, c->d_surv.p, c->d_nsurv.p, c->num_cus, c->stream));
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    if (kernel_ms) *kernel_ms = ms;
    return VOSS_OK;
}

int voss_mr_batch_device_noalloc(voss_ctx* c, const uint64_t* d_n, size_t count,
                                 uint8_t* d_verdict, double* kernel_ms) {
    VOSS_ADMIT(c);
    if (!c) return fail(VOSS_EINVAL, "ctx is NULL");
    if (count && (!d_n || !d_verdict)) return fail(VOSS_EINVAL, "NULL buffer");
    if (count > kMrNoAllocMax)
        return fail(VOSS_EINVAL, "count beyond VOSS_MR_NOALLOC_MAX (4194304) candidates");
    if (!c->mr_scratch) return fail(VOSS_EDEVICE, "no scratch address for this device");
    int rc;
    if ((rc = set_device(c))) return rc;
    if (count == 0) {
        if (kernel_ms) *kernel_ms = 0;
        return VOSS_OK;
    }
    // the fixed array instead of c->d_surv / c->d_nsurv: launch_mr clears the
    // counters with a fill kernel, and the lists are written before read
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    HIPCHK(launch_mr(d_n, count, d_verdict, c->mr_scratch, c->mr_scratch + 2 * kMrNoAllocMax,
                     c->num_cus, c->stream));
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    if (kernel_ms) *kernel_ms = ms;
    return VOSS_OK;
}

int voss_mr_batch(voss_ctx* c, const uint64_t* n, size_t count, uint8_t* verdict) {
    VOSS_ADMIT(c);
    if (!c) return fail(VOSS_EINVAL, "ctx is NULL");
    if (count && (!n || !verdict)) return fail(VOSS_EINVAL, "NULL buffer");
    int rc;
    if ((rc = set_device(c))) return rc;
    if (count == 0) return VOSS_OK;
    HIPCHK(c->d_mr_n.reserve(count));
    HIPCHK(c->d_mr_v.reserve(count));
    HIPCHK(hipMemcpyAsync(c->d_mr_n.p, n, count * 8, hipMemcpyHostToDevice, c->stream));
    if ((rc = voss_mr_batch_device(c, c->d_mr_n.p, count, c->d_mr_v.p, nullptr))) return rc;
    HIPCHK(hipMemcpy(verdict, c->d_mr_v.p, count, hipMemcpyDeviceToHost));
    return VOSS_OK;
}

int voss_verify_gaps(voss_ctx* c, const uint64_t* p_after, const int32_t* gap, size_t count,
                     uint8_t* ok) {
    VOSS_ADMIT(c);
    if (!c) return fail(VOSS_EINVAL, "ctx is NULL");
    if (count && (!p_after || !gap || !ok)) return fail(VOSS_EINVAL, "NULL buffer");
    if (count == 0) return VOSS_OK;
    // both ends: GPU Miller-Rabin (verify_gap :61-62)
    std::vector<uint64_t> ends(2 * count);
    for (size_t i = 0; i < count; i++) {
        ends[2 * i] = p_after[i] - (uint64_t)(int64_t)gap[i];
        ends[2 * i + 1] = p_after[i];
    }
    std::vector<uint8_t> v(2 * count);
    int rc = voss_mr_batch(c, ends.data(), ends.size(), v.data());
    if (rc) return rc;
    // interior: no prime strictly between the ends (:63-92), decided on the
    // device for every row (verify_interior_kernel: one wave per gap tests
    // its interior numbers with the same Miller-Rabin; claimed gaps beyond
    // p_after fail, as in the reference)
    for (size_t i = 0; i < count; i++) ok[i] = (uint8_t)(v[2 * i] && v[2 * i + 1]);
    if (count >= 0xFFFFFFFFull) return fail(VOSS_EINVAL, "count must be below 2^32");
    if ((rc = set_device(c))) return rc;
    HIPCHK(c->d_vf_pa.reserve(count));
    HIPCHK(c->d_vf_gap.reserve(count));
    HIPCHK(c->d_vf_ok.reserve(count));
    HIPCHK(hipMemcpyAsync(c->d_vf_pa.p, p_after, count * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->d_vf_gap.p, gap, count * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->d_vf_ok.p, ok, count, hipMemcpyHostToDevice, c->stream));
    HIPCHK(launch_verify_interior(c->d_vf_pa.p, c->d_vf_gap.p, c->d_vf_ok.p, (uint32_t)count,
                                  c->stream));
    HIPCHK(hipMemcpyAsync(ok, c->d_vf_ok.p, count, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return VOSS_OK;
}

int voss_device_alloc(voss_ctx* c, size_t bytes, void** dptr) {
    VOSS_ADMIT(c);
    if (!c || !dptr) return fail(VOSS_EINVAL, "NULL argument");
    int rc = set_device(c);
    if (rc) return rc;
    HIPCHK(hipMalloc(dptr, bytes));
    return VOSS_OK;
}

int voss_device_free(voss_ctx* c, void* dptr) {
    VOSS_ADMIT(c);
    if (!c) return fail(VOSS_EINVAL, "ctx is NULL");
    int rc = set_device(c);
    if (rc) return rc;
    HIPCHK(hipFree(dptr));
    return VOSS_OK;
}

int voss_memcpy_h2d(voss_ctx* c, void* dst, const void* src, size_t bytes) {
    VOSS_ADMIT(c);
    if (!c) return fail(VOSS_EINVAL, "ctx is NULL");
    int rc = set_device(c);
    if (rc) return rc;
    HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return VOSS_OK;
}

int voss_memcpy_d2h(voss_ctx* c, void* dst, const void* src, size_t bytes) {
    VOSS_ADMIT(c);
    if (!c) return fail(VOSS_EINVAL, "ctx is NULL");
    int rc = set_device(c);
    if (rc) return rc;
    HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return VOSS_OK;
}

}  // extern "C"
