// sieve_device.h -- device code of the wheel-30 window sieve and of the
// large-prime bucket fill, shared by the launch-per-plan kernels (sieve.hip)
// and the persistent serving loop (serve_kernel.hip).
//
// One workgroup sieves one LDS window at a time (sieve_window):
//   phase 0  the window's bitmap is initialised from 4 presieve pattern
//            tables for primes 7..37, one AND per table per word;
//   phase 1  the window's bucket hits (primes >= bucket_limit, precomputed in
//            HBM) are applied, then every other base prime p with
//            p*p <= window end crosses off its 8 residue classes with
//            ds_and_b32.  Primes below tier_a_end are marked by a whole wave
//            (8 lanes per class, stride 8p); larger ones by one lane each.
//            Waves pull work from an LDS counter, smallest primes first;
//   phase 2  each thread walks its contiguous run of words in increasing
//            order: popcounts give the per-residue and mod-4 counts; one
//            loop iteration per word, its four wheel bytes unrolled: a
//            nonzero byte's first prime pairs with the previous byte's last
//            one (one LDS atomic into the histograms) and the pairs inside
//            the byte are counted by byte value; gaps that cross thread
//            runs are joined after a barrier.
// The histograms stay in LDS across windows of one job and are merged into
// the job's HBM accumulators by sieve_flush.  The bitmap never leaves LDS.
#pragma once
#include <hip/hip_runtime.h>

#include "fp_recip.h"
#include <type_traits>
#include "sieve_launch.h"
#include "wheel.h"

namespace voss {

// Residue r (0..29) -> wheel bit of r (-1 if r is not coprime to 30); the
// smallest wheel bit o with W30[o] >= r (8 if none); the largest o with
// W30[o] <= r (-1 if none).  Initialised constants: no upload per device.
#define VOSS_RESIDX {-1, 0, -1, -1, -1, -1, -1, 1, -1, -1, -1, 2, -1, 3, -1, \
                     -1, -1, 4, -1, 5, -1, -1, -1, 6, -1, -1, -1, -1, -1, 7}
#define VOSS_FIRSTBIT {0, 0, 1, 1, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 4, \
                       4, 4, 4, 5, 5, 6, 6, 6, 6, 7, 7, 7, 7, 7, 7}
#define VOSS_LASTBIT {-1, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 3, 3, \
                      3, 3, 4, 4, 5, 5, 5, 5, 6, 6, 6, 6, 6, 6, 7}
static __constant__ int8_t c_residx[30] = VOSS_RESIDX;
static __constant__ int8_t c_firstbit[30] = VOSS_FIRSTBIT;
static __constant__ int8_t c_lastbit[30] = VOSS_LASTBIT;

namespace tables_check {
constexpr int8_t kResIdx[30] = VOSS_RESIDX;
constexpr int8_t kFirstBit[30] = VOSS_FIRSTBIT;
constexpr int8_t kLastBit[30] = VOSS_LASTBIT;
constexpr bool ok() {
    for (int r = 0; r < 30; r++) {
        int idx = -1, fb = 8, lb = -1;
        for (int o = 0; o < 8; o++) {
            if (kW30[o] == r) idx = o;
            if (kW30[o] <= r) lb = o;
        }
        for (int o = 7; o >= 0; o--)
            if (kW30[o] >= r) fb = o;
        if (kResIdx[r] != idx || kFirstBit[r] != fb || kLastBit[r] != lb) return false;
    }
    return true;
}
static_assert(ok(), "wheel residue tables");
}  // namespace tables_check
#undef VOSS_RESIDX
#undef VOSS_FIRSTBIT
#undef VOSS_LASTBIT



// W30[o] as eight bytes of one 64-bit constant, and W30_TO_Q6_IDX
// (src/voss_v7_live.cu:30) as the bits of 0xD4: no table loads in the
// per-prime path.
constexpr unsigned long long kW30Bytes = 0x1D1713110D0B0701ull;
__device__ __forceinline__ uint32_t w30_of(uint32_t o) {
    return (uint32_t)(kW30Bytes >> (8u * o)) & 0xFFu;
}
__device__ __forceinline__ uint32_t q6row_of(uint32_t o) { return (0xD4u >> o) & 1u; }

// Bit 8j+7 set iff byte j of w is nonzero.
__device__ __forceinline__ uint32_t nonzero_bytes(uint32_t w) {
    return (((w & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | w) & 0x80808080u;
}

// x mod p for any 64-bit x: m = floor(2^64 / p) (Barrett, one correction).
__device__ __forceinline__ uint32_t mod_u64(uint64_t x, uint32_t p,
                                            uint64_t m) {
    uint64_t q = __umul64hi(x, m);
    uint64_t r = x - q * (uint64_t)p;
    return (uint32_t)(r >= p ? r - p : r);
}

// X mod p in double precision for many p and one X (a window or tile
// start; start_mod): xd = X rounded to double, delta = X - xd.  Usable (fp)
// when X < 2^63, every p is below 2^30 and every X/p below 2^47.
struct ModStart {
    uint64_t X;
    double xd;
    int32_t delta;
    bool fp;
};

__device__ __forceinline__ ModStart mod_start(uint64_t X, uint32_t pmin, uint32_t pmax) {
    ModStart t;
    t.X = X;
    t.fp = X < (1ull << 63) && pmax < (1u << 30) && (X >> 47) < pmin;
    t.xd = (double)X;
    t.delta = t.fp ? (int32_t)(int64_t)(X - (uint64_t)t.xd) : 0;
    return t;
}

// X mod p for t.fp starts without 64-bit integer multiplies: 1/p from
// recip_newton (relative error below 2^-48), the quotient estimate
// q = floor(xd*inv) (xd*inv is within 2^-47.9 relative of X/p < 2^47, so
// within 1/2 of it, and q within 1 of the quotient), then
// X - q*p = fma(-q, p, xd) + delta, which is exact (an integer below 2^31
// in magnitude), brought into [0, p) by one correction each way.  (A second
// Newton step cost 1.4 % of the 4096-interval batch's bucket fill.)
__device__ __forceinline__ uint32_t start_mod(const ModStart& t, uint32_t p) {
    const double pd = (double)p;
    const double inv = recip_newton(pd);
    const double q = __builtin_floor(t.xd * inv);
    int32_t r = (int32_t)__builtin_fma(-q, pd, t.xd) + t.delta;
    r = r < 0 ? r + (int32_t)p : r;
    r = r >= (int32_t)p ? r - (int32_t)p : r;
    return (uint32_t)r;
}

__device__ __forceinline__ void clear_bit(uint32_t* bm, uint32_t b,
                                          uint32_t bit) {
    atomicAnd(&bm[b >> 2], ~(1u << (((b & 3u) << 3) | bit)));
}

__device__ __forceinline__ uint32_t clear_mask(uint32_t b, uint32_t bit) {
    return ~(1u << (((b & 3u) << 3) | bit));
}

// ds_and_b32 at LDS byte address a (a multiple of 4) of the bitmap.
__device__ __forceinline__ void and_at(uint32_t* bm, uint32_t a, uint32_t m) {
    atomicAnd(reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(bm) + a), m);
}

// Crosses bytes b, b+p, b+2p, ... < len of one residue class (len is a
// multiple of 4, so byte b < len iff its word address b & ~3 < len).  Four
// consecutive marks are one step: each keeps its own word address, which
// advances by 4p (a multiple of 4, so it stays word-aligned and exact), and
// its own mask.  p is odd, so from one mark to the next the marked byte
// moves by p mod 4 within the word: the next mask is the previous one
// rotated by 8 * (p mod 4) bits.
__device__ __forceinline__ void cross_class(uint32_t* bm, uint32_t b, uint32_t p,
                                            uint32_t bit, uint32_t len) {
    const uint32_t rot = 32u - 8u * (p & 3u);  // rotate right by this = left by 8*(p%4)
    const uint32_t m0 = clear_mask(b, bit);
    const uint32_t m1 = __builtin_amdgcn_alignbit(m0, m0, rot);
    const uint32_t m2 = __builtin_amdgcn_alignbit(m1, m1, rot);
    const uint32_t m3 = __builtin_amdgcn_alignbit(m2, m2, rot);
    uint32_t a0 = b & ~3u, a1 = (b + p) & ~3u, a2 = (b + 2u * p) & ~3u, a3 = (b + 3u * p) & ~3u;
    const uint32_t p4 = 4u * p;
    for (; a3 < len; a0 += p4, a1 += p4, a2 += p4, a3 += p4) {
        and_at(bm, a0, m0);
        and_at(bm, a1, m1);
        and_at(bm, a2, m2);
        and_at(bm, a3, m3);
    }
    if (a0 < len) and_at(bm, a0, m0);
    if (a1 < len) and_at(bm, a1, m1);
    if (a2 < len) and_at(bm, a2, m2);
}

// Position of the first multiple of p in residue class j inside the window
// starting at byte w0 (relative byte index, may be >= len).  The multiple
// p*1 (the prime itself) is skipped; smaller composite multiples p*m with
// 7 <= m < p may be crossed, which changes nothing.
__device__ __forceinline__ uint32_t first_pos(uint32_t p, uint32_t r,
                                              int j, uint64_t w0) {
    // p*W30[j] fits 32 bits for every base prime of numbers below 2^54
    const uint32_t off = p < 148102320u ? (p * w30_of((uint32_t)j)) / 30u
                                        : (uint32_t)(((uint64_t)p * w30_of((uint32_t)j)) / 30u);
    uint32_t pos = off >= r ? off - r : off + p - r;
    if (j == 0 && w0 <= (uint64_t)off) pos += p;
    return pos;
}

// The sieve's per-(prime, window) setup without 32-bit multiplies or
// divisions per class (those are quarter-rate on the VALU):
//   floor(p*k/30) = pq*k + floor(pr*k/30)   (pq = p / 30, pr = p % 30,
//   k = W30[j]), and floor(pr*k/30) = (pr*k*2185) >> 16 exactly for
//   pr, k < 30 (24-bit multiply).  pqk = pq*k.
__device__ __forceinline__ uint32_t class_off(uint32_t pqk, uint32_t pr, uint32_t k) {
    return pqk + (__umul24(pr, k * 2185u) >> 16);
}

// first_pos from the class offset off = floor(p*W30[j]/30) < p and
// r = w0 mod p < p: (off - r) mod p is the smaller of off - r and
// off - r + p as unsigned 32-bit values.
__device__ __forceinline__ uint32_t first_pos_off(uint32_t off, uint32_t p, uint32_t r,
                                                  bool cls0, uint64_t w0) {
    const uint32_t d = off - r;
    uint32_t pos = min(d, d + p);
    if (cls0 && w0 <= (uint64_t)off) pos += p;
    return pos;
}

__device__ __forceinline__ uint32_t prime_mod30(uint32_t perm) { return (perm >> 24) & 31u; }

// wheel bit of residue r (r coprime to 30), four bits per residue
constexpr unsigned long long res_nibbles(int half) {
    unsigned long long v = 0;
    for (int o = 0; o < 8; o++) {
        const int r = (int)((kW30Bytes >> (8 * o)) & 0xFF);
        if ((r >> 4) == half) v |= (unsigned long long)o << (4 * (r & 15));
    }
    return v;
}
constexpr unsigned long long kResLo = res_nibbles(0), kResHi = res_nibbles(1);
constexpr uint32_t kCoprime30 = (1u << 1) | (1u << 7) | (1u << 11) | (1u << 13) | (1u << 17) |
                                (1u << 19) | (1u << 23) | (1u << 29);
__device__ __forceinline__ uint32_t wheel_bit_of_residue(uint32_t r) {
    return (uint32_t)(((r < 16 ? kResLo : kResHi) >> (4u * (r & 15u))) & 7u);
}


// Marks m*p for every m >= 2 coprime to 30 with m*p in the window
// [X, X + 30*len): from r = X mod p the first multiple, then
// m -> m+1 steps (p/30 bytes and p%30 in the residue), keeping those
// coprime to 30.  For p >= 30*len/4 this visits at most 5 multiples, fewer
// than the 8 residue-class starts of the class-by-class marking (from
// 30*len/7, up to 7 multiples, it was slower: 1e13 interval sieve 178.0
// against 172.2 ms).
__device__ __forceinline__ void walk_prime(uint32_t* bm, uint32_t p, uint32_t r, uint64_t X,
                                           uint32_t len) {
    // r = X mod p, X = 30*w0 the window's first number
    const uint32_t x = X < 2ull * p ? (uint32_t)(2ull * p - X) : (r ? p - r : 0u);
    uint32_t b = x / 30u, rr = x - 30u * b;
    const uint32_t pd = p / 30u, pm = p - 30u * pd;
    while (b < len) {
        if ((kCoprime30 >> rr) & 1u) clear_bit(bm, b, wheel_bit_of_residue(rr));
        b += pd;
        rr += pm;
        if (rr >= 30u) {
            rr -= 30u;
            b++;
        }
    }
}

// Sum of x over the wave, valid in lane 63: a row-wise inclusive scan
// (DPP row_shr 1, 2, 4, 8), then the row totals carried by row_bcast 15
// and 31 (VALU data-parallel moves, no LDS round trips).
__device__ __forceinline__ uint32_t wave_sum63(uint32_t x) {
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xf, 0xf, false);
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xf, 0xf, false);
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xf, 0xf, false);
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xf, 0xf, false);
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xa, 0xf, false);
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xc, 0xf, false);
    return x;
}

// Thresholds of the merit rule for the even gaps 102..500 (bins 51..250)
// live in LDS: lgt[h - 51].
constexpr int kLgtFirst = 51, kLgtCount = 200;

struct PairCtx {
    uint32_t* hist;
    const unsigned long long* lgt;
    JobAcc* acc;                  // this window's job
    LargeRaw* large;
    unsigned long long* large_n;
    unsigned long long large_cap;
    unsigned long long base_num;
    uint32_t job;
    uint32_t g_maybe;  // below it no gap of this window can be large
};

// A consecutive pair inside a window with gap g >= 100 (pb = offset of
// p_before from base_num).  Gaps below 100 only touch the q=30 row
// (mq[8*(g/2) + row]); the q=6 rows and the main histogram follow at flush.
// Inlined on purpose: a call would have to spill the context to scratch.
__device__ __forceinline__ void pair_wide(const PairCtx& c, uint32_t pb, uint32_t g) {
    if (g >= 2u * kHistBins) {
        atomicOr(&c.acc->error, 1ull);
        return;
    }
    atomicAdd(&c.hist[g >> 1], 1u);
    if (g >= c.g_maybe) {
        const unsigned long long p_before = c.base_num + pb;
        if (g >= (uint32_t)kLargeAlways || p_before <= c.lgt[(g >> 1) - kLgtFirst]) {
            const unsigned long long slot = atomicAdd(c.large_n, 1ull);
            atomicAdd(&c.acc->large_count, 1ull);
            if (slot < c.large_cap) {
                LargeRaw r;
                r.p_before = p_before;
                r.gap = g;
                r.job = c.job;
                c.large[slot] = r;
            }
        }
    }
}

// The large-gap part of pair_wide for a pair whose histogram count is
// already taken (g >= c.g_maybe, so g >= 102): a gap beyond the device
// histogram sets the error bit; a large gap by the reference's rule
// (g >= 500, or merit >= 10 via the threshold table) becomes a candidate.
__device__ __forceinline__ void pair_large(const PairCtx& c, uint32_t pb, uint32_t g) {
    if (g >= 2u * kHistBins) {
        atomicOr(&c.acc->error, 1ull);
        return;
    }
    const unsigned long long p_before = c.base_num + pb;
    if (g >= (uint32_t)kLargeAlways || p_before <= c.lgt[(g >> 1) - kLgtFirst]) {
        const unsigned long long slot = atomicAdd(c.large_n, 1ull);
        atomicAdd(&c.acc->large_count, 1ull);
        if (slot < c.large_cap) {
            LargeRaw r;
            r.p_before = p_before;
            r.gap = g;
            r.job = c.job;
            c.large[slot] = r;
        }
    }
}

// LDS of one sieving workgroup.  A 128 KiB window keeps the presieve tables
// in LDS (one workgroup per CU); smaller windows read them from global
// memory (12.6 KB, L1 resident) so that two 64 KiB or four 32 KiB
// workgroups fit in a CU's LDS.
template <int WIN, int BLOCK, bool EMIT>
struct alignas(16) SieveShared {
    static constexpr bool kPatLds = WIN > 65536;
    uint32_t bm[WIN / 4];
    uint32_t hist[kHistBins];
    // q=30 rows interleaved: mq[8*(g/2) + row] for gaps g < 100; bin 0 takes
    // the walk's gap-0 self-pairs (dropped at the flush); bins 50.. unused.
    // The walk's odd lanes count into mq2, so two lanes of a wave that hit
    // the same bin more often hit different words (summed at the flush)
    uint32_t mq[64 * 8];
    uint32_t mq2[50 * 8];
    // walk: how many nonzero wheel bytes of each value were seen (their
    // internal prime pairs go into mq at the flush)
    uint32_t bytes[EMIT ? 1 : 256];
    // walk: per wheel byte value v, W30 of its lowest set bit (bits 0-4),
    // W30 of its highest (5-9), 4 * the highest bit (10-14), and bit 15 set
    // when v holds two primes or more
    uint16_t blut[EMIT ? 1 : 256];
    uint32_t pat[kPatLds ? kPresieveWordsPadded : 1];
    uint32_t tfirst[BLOCK];
    uint32_t tlast[BLOCK];
    unsigned long long cnt[10];
    unsigned long long lgt[kLgtCount];
    uint32_t task;
    uint32_t gmaybe;
    uint32_t next;
    uint32_t patbase[kPresieveGroups];
    uint32_t desc[sizeof(SegDesc) / 4];
};
static_assert(sizeof(SegDesc) % 4 == 0, "descriptor dwords");

// Where one window's results go.
struct WinIO {
    JobAcc* acc;                  // the window's job
    LargeRaw* large;              // large-gap candidate list
    unsigned long long* large_n;  // its length (one atomic slot per record)
    unsigned long long large_cap;
    uint32_t job;
    const uint32_t* bent;         // the window's bucket entries (16-byte aligned), or null
    uint32_t nbent;               // how many (clamped to the bucket's room)
    SegRec* rec;                  // the window's record
    unsigned long long emit_at;   // EMIT: first output slot of the window
};

// Clears the histograms and loads the tables every window reads.
template <int WIN, int BLOCK, bool EMIT>
__device__ __forceinline__ void sieve_init_lds(const SieveArgs& a, SieveShared<WIN, BLOCK, EMIT>& S) {
    const int tid = threadIdx.x;
    for (int i = tid; i < kHistBins; i += BLOCK) S.hist[i] = 0;
    for (int i = tid; i < 8 * 64; i += BLOCK) S.mq[i] = 0;
    for (int i = tid; i < 8 * 50; i += BLOCK) S.mq2[i] = 0;
    if constexpr (!EMIT) {
        for (int i = tid; i < 256; i += BLOCK) {
            S.bytes[i] = 0;
            const uint32_t v = (uint32_t)i;
            const uint32_t lo = (uint32_t)__builtin_ctz(v | 0x100u), hi = 31u - (uint32_t)__builtin_clz(v | 1u);
            S.blut[i] = v ? (uint16_t)(w30_of(lo) | (w30_of(hi) << 5) | ((4u * hi) << 10) |
                                       ((v & (v - 1u)) ? 0x8000u : 0u))
                          : (uint16_t)0;
        }
    }
    if constexpr (SieveShared<WIN, BLOCK, EMIT>::kPatLds)
        for (int i = tid; i < kPresieveWordsPadded; i += BLOCK) S.pat[i] = a.pat[i];
    for (int i = tid; i < kLgtCount; i += BLOCK) S.lgt[i] = a.lg_thresh[kLgtFirst + i];
}

// Presieve table offsets of the window starting at byte w0 (threads below
// kPresieveGroups).
template <int WIN, int BLOCK, bool EMIT>
__device__ __forceinline__ void sieve_set_patbase(const SieveArgs& a, SieveShared<WIN, BLOCK, EMIT>& S,
                                                  uint64_t w0, int t) {
    if (t < kPresieveGroups) S.patbase[t] = (uint32_t)((w0 >> 2) % a.pat_len[t]);
}

// Merges the workgroup's histograms into acc (u64 atomics, so the sums do
// not depend on scheduling) and clears them.  Every thread calls it.
template <int WIN, int BLOCK, bool EMIT>
__device__ __forceinline__ void sieve_flush(SieveShared<WIN, BLOCK, EMIT>& S, JobAcc* acc) {
    const int tid = threadIdx.x;
    __syncthreads();
    if constexpr (!EMIT) {
        // the consecutive set bits o1 < o2 of a byte value are a prime
        // pair with gap W30[o2] - W30[o1] and p_before in row o1
        for (int v = tid; v < 256; v += BLOCK) {
            const uint32_t c = S.bytes[v];
            if (!c) continue;
            S.bytes[v] = 0;
            uint32_t x = (uint32_t)v, o1 = __builtin_ctz(x);
            x &= x - 1u;
            while (x) {
                const uint32_t o2 = __builtin_ctz(x);
                x &= x - 1u;
                atomicAdd(&S.mq[((w30_of(o2) - w30_of(o1)) >> 1) * 8u + o1], c);
                o1 = o2;
            }
        }
        __syncthreads();
    }
    for (int i = tid; i < kHistBins; i += BLOCK) {
        uint32_t v = S.hist[i];
        if (v) {
            atomicAdd(&acc->hist[i], (unsigned long long)v);
            S.hist[i] = 0;
        }
    }
    for (int i = tid; i < 8 * 64; i += BLOCK) {
        const uint32_t v = S.mq[i] + (i < 8 * 50 ? S.mq2[i] : 0u);
        S.mq[i] = 0;
        if (i < 8 * 50) S.mq2[i] = 0;
        if (v && i >= 8 && i < 8 * 50) {  // bins 1..49: gaps 2..98
            const uint32_t row = (uint32_t)i & 7u, h = (uint32_t)i >> 3;
            const unsigned long long v64 = v;
            atomicAdd(&acc->modq[(2u + row) * kModqBins + h], v64);
            atomicAdd(&acc->modq[q6row_of(row) * kModqBins + h], v64);
            atomicAdd(&acc->hist[h], v64);
        }
    }
    __syncthreads();
}

// Sieves window D into S, adds its counts and pairs into io, writes its
// record.  S.patbase must hold D's presieve offsets.  The last wave calls
// prefetch() at the start of marking (whose work is pulled from a counter,
// so the wave that runs it takes fewer tasks): the caller fetches what the
// next window needs there.  Every thread calls it; the histograms stay in S
// until sieve_flush.
// COHERENT_BUCKETS: the bucket entries were written by other workgroups of
// the same running kernel and are read with device-coherent loads.
template <int WIN, int BLOCK, bool EMIT, bool COHERENT_BUCKETS = false, typename Prefetch>
__device__ __forceinline__ void sieve_window(const SieveArgs& a, SieveShared<WIN, BLOCK, EMIT>& S,
                                             const SegDesc& D, const WinIO& io, Prefetch&& prefetch) {
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const uint32_t* const pat = SieveShared<WIN, BLOCK, EMIT>::kPatLds ? S.pat : a.pat;
    uint32_t* const bm = S.bm;
    const uint32_t L4 = D.len >> 2;
    if (tid == 0) {
        S.task = 0;
        S.gmaybe = (uint32_t)kLargeAlways;
    }
    if (tid < 10) S.cnt[tid] = 0;
    __syncthreads();

    // lgt grows with the gap: the smallest gap whose threshold reaches
    // the window start bounds the gaps that may be large in this window
    if (tid < kLgtCount && S.lgt[tid] >= 30ull * D.w0)
        atomicMin(&S.gmaybe, 2u * (uint32_t)(kLgtFirst + tid));

    // ---- phase 0: presieve patterns (primes 7..37) from the tables
    {
        // each thread fills word pairs (2t, 2t+1), (2t+2B, 2t+2B+1), ...:
        // a pair reads two consecutive pattern words per table (the
        // padding word makes the second read safe at the table's end)
        uint32_t idx[kPresieveGroups];
#pragma unroll
        for (int g = 0; g < kPresieveGroups; g++)
            idx[g] = (S.patbase[g] + 2u * (uint32_t)tid) % a.pat_len[g];
        for (uint32_t i = 2u * tid; i < L4; i += 2u * BLOCK) {
            uint32_t w0 = 0xFFFFFFFFu, w1 = 0xFFFFFFFFu;
#pragma unroll
            for (int g = 0; g < kPresieveGroups; g++) {
                const uint32_t* t = pat + a.pat_off[g] + idx[g];
                w0 &= t[0];
                w1 &= t[1];
                const uint32_t v = idx[g] + a.pat_step[g];
                idx[g] = v >= a.pat_len[g] ? v - a.pat_len[g] : v;
            }
            if (i == 0 && D.w0 == 0) w0 = (w0 | a.restore0) & ~1u;
            *reinterpret_cast<uint2*>(&bm[i]) = make_uint2(w0, w1);
        }
    }
    __syncthreads();

    // ---- phase 1: base primes 79 <= p, p*p <= window end
    // (the presieve above has read S.patbase, and D is in registers)
    if (tid >= BLOCK - 64) prefetch();
    {
        const uint32_t np = D.nprimes;
        const uint32_t a_end = np < a.tier_a_end ? np : a.tier_a_end;
        const uint32_t nA = a_end > a.tier_a_begin ? a_end - a.tier_a_begin : 0;
        // tier B1: primes below len/4 (below len in a partial window), 64
        // per task, a loop per class; tier B2: primes >= len/4 (a few hits
        // each), 256 per task with four record loads in flight per lane, so
        // the loads' latency is paid once per task, and straight-line
        // marking: in a full window [len/4, len/3), [len/3, len/2),
        // [len/2, len) and [len, ...) have at most 4, 3, 2 and 1 hits per
        // residue class, each range its own tasks
        const uint32_t b2 = max(a.tier_a_end, min(np, D.ibig));
        const bool full = D.len == (uint32_t)WIN;
        const uint32_t c2 = full ? max(a.tier_a_end, min(b2, a.ib_frac[2])) : b2;
        const uint32_t c3 = full ? max(a.tier_a_end, min(c2, a.ib_frac[1])) : b2;
        const uint32_t c4 = full ? max(a.tier_a_end, min(c3, a.ib_frac[0])) : b2;
        const uint32_t nB1 = c4 > a.tier_a_end ? (c4 - a.tier_a_end + 63) / 64 : 0;
        const uint32_t n4 = (c3 - c4 + 255) / 256, n3 = (c2 - c3 + 255) / 256;
        const uint32_t n2 = (b2 - c2 + 255) / 256;
        const uint32_t nB2 = np > b2 ? (np - b2 + 255) / 256 : 0;
        const uint32_t ntask = nA + nB1 + n4 + n3 + n2 + nB2;
        const uint32_t len = D.len;
        if (io.bent) {
            // large primes: their hits in this window, precomputed.
            // Buckets start 16-byte aligned and their room is a multiple
            // of 4 entries, so every lane reads 4 entries with one load.
            const uint32_t ne = io.nbent;
            const uint4* ent4 = reinterpret_cast<const uint4*>(io.bent);
            for (uint32_t k = tid; 4u * k < ne; k += BLOCK) {
                uint4 e4;
                if constexpr (COHERENT_BUCKETS) {
                    unsigned long long* q = const_cast<unsigned long long*>(
                        reinterpret_cast<const unsigned long long*>(ent4 + k));
                    const unsigned long long x0 =
                        __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    const unsigned long long x1 =
                        __hip_atomic_load(q + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    e4 = make_uint4((uint32_t)x0, (uint32_t)(x0 >> 32), (uint32_t)x1,
                                    (uint32_t)(x1 >> 32));
                } else {
                    e4 = ent4[k];
                }
                const uint32_t ev[4] = {e4.x, e4.y, e4.z, e4.w};
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    if (4u * k + q < ne) {
                        const uint32_t b = ev[q] >> 3;
                        atomicAnd(&bm[b >> 2], ~(1u << (((b & 3u) << 3) | (ev[q] & 7u))));
                    }
                }
            }
        }
        while (true) {
            uint32_t t = 0;
            if (lane == 0) t = atomicAdd(&S.task, 1u);
            t = __shfl(t, 0);
            if (t >= ntask) break;
            if (t < nA) {
                const uint32_t pi = a.tier_a_begin + t;
                const PrimeRec pr = a.primes[pi];
                const uint32_t p = pr.p;
                const uint32_t r = mod_u64(D.w0, p, pr.magic);
                const int j = lane >> 3;
                const uint32_t s = lane & 7;
                const uint32_t bit = (pr.perm >> (3 * j)) & 7u;
                // p < tier_a_limit: 24-bit products
                const uint32_t k = w30_of((uint32_t)j);
                const uint32_t off = class_off(__umul24(p / 30u, k), prime_mod30(pr.perm), k);
                const uint32_t b = first_pos_off(off, p, r, j == 0, D.w0) + __umul24(s, p);
                // stride 8p = 0 mod 4: one mask, and the word address
                // b & ~3 advances exactly by the stride
                const uint32_t stride = 8u * p;
                const uint32_t m = clear_mask(b, bit);
                uint32_t a4 = b & ~3u;
                for (; a4 + 3u * stride < len; a4 += 4u * stride) {
                    and_at(bm, a4, m);
                    and_at(bm, a4 + stride, m);
                    and_at(bm, a4 + 2u * stride, m);
                    and_at(bm, a4 + 3u * stride, m);
                }
                for (; a4 < len; a4 += stride) and_at(bm, a4, m);
            } else if (t < nA + nB1) {
                const uint32_t pi = a.tier_a_end + (t - nA) * 64 + lane;
                if (pi < c4) {
                    const PrimeRec pr = a.primes[pi];
                    const uint32_t p = pr.p;
                    const uint32_t r = mod_u64(D.w0, p, pr.magic);
                    const uint32_t pq = p / 30u, pm = prime_mod30(pr.perm);
#pragma unroll 1
                    for (int j = 0; j < 8; j++) {
                        // p < len: 24-bit products
                        const uint32_t k = w30_of((uint32_t)j);
                        const uint32_t off = class_off(__umul24(pq, k), pm, k);
                        cross_class(bm, first_pos_off(off, p, r, j == 0, D.w0), p,
                                    (pr.perm >> (3 * j)) & 7u, len);
                    }
                }
            } else {
                // primes [first, end) of the range with at most HITS hits
                // per class; task tt of the range
                auto mark_few = [&](auto hits_c, uint32_t first, uint32_t end, uint32_t tt) {
                    constexpr int HITS = decltype(hits_c)::value;
                    const uint32_t pi0 = first + tt * 256 + lane;
                    PrimeRec pr[4];
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        const uint32_t pi = pi0 + 64u * q;
                        pr[q] = a.primes[pi < end ? pi : end - 1u];
                    }
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        if (pi0 + 64u * q >= end) continue;
                        const uint32_t p = pr[q].p;
                        if (HITS == 1 && 4ull * p >= 30ull * len) {
                            // (start_mod's double-precision remainder measured
                            // slower here than the Barrett reduction)
                            const uint64_t X = 30ull * D.w0;
                            walk_prime(bm, p, mod_u64(X, p, pr[q].magic), X, len);
                        } else {
                            // at most HITS multiples per class in the window,
                            // p bytes apart; pq*W30[j] by additions (W30
                            // steps 6,4,2,4,2,4,6)
                            const uint32_t r = mod_u64(D.w0, p, pr[q].magic);
                            const uint32_t pq = p / 30u, pm = prime_mod30(pr[q].perm);
                            uint32_t pq2 = pq + pq, pq4 = pq2 + pq2;
                            // opaque to the compiler, which would otherwise
                            // fold the additions back into multiplies
                            asm volatile("" : "+v"(pq2), "+v"(pq4));
                            const uint32_t pq6 = pq4 + pq2;
                            uint32_t pqk = pq;
#pragma unroll
                            for (int j = 0; j < 8; j++) {
                                const uint32_t k = w30_of((uint32_t)j);
                                if (j > 0) {
                                    const uint32_t dk = k - w30_of((uint32_t)j - 1u);
                                    pqk += dk == 6u ? pq6 : dk == 4u ? pq4 : pq2;
                                }
                                const uint32_t b =
                                    first_pos_off(class_off(pqk, pm, k), p, r, j == 0, D.w0);
                                const uint32_t bit = (pr[q].perm >> (3 * j)) & 7u;
                                if (b < len) clear_bit(bm, b, bit);
                                if (HITS >= 2 && b + p < len) clear_bit(bm, b + p, bit);
                                if (HITS >= 3 && b + 2u * p < len) clear_bit(bm, b + 2u * p, bit);
                                if (HITS >= 4 && b + 3u * p < len) clear_bit(bm, b + 3u * p, bit);
                            }
                        }
                    }
                };
                uint32_t u = t - nA - nB1;
                if (u < n4) mark_few(std::integral_constant<int, 4>{}, c4, c3, u);
                else if ((u -= n4) < n3) mark_few(std::integral_constant<int, 3>{}, c3, c2, u);
                else if ((u -= n3) < n2) mark_few(std::integral_constant<int, 2>{}, c2, b2, u);
                else mark_few(std::integral_constant<int, 1>{}, b2, np, u - n2);
            }
        }
    }
    __syncthreads();

    // ---- phase 2: mask to [n_lo, n_hi], count, walk primes in order
    const unsigned long long base_num = 30ull * D.w0;
    int bitlo = 0, bithi = (int)(8u * D.len) - 1;
    if (D.n_lo > base_num) {
        uint32_t rel = (uint32_t)(D.n_lo - base_num);
        bitlo = (int)((rel / 30u) * 8u) + c_firstbit[rel % 30u];
    }
    if (D.n_hi < base_num + 30ull * D.len - 1ull) {
        uint32_t rel = (uint32_t)(D.n_hi - base_num);
        bithi = (int)((rel / 30u) * 8u) + c_lastbit[rel % 30u];
    }
    const PairCtx pc{S.hist, S.lgt, io.acc, io.large, io.large_n, io.large_cap,
                     base_num, io.job, S.gmaybe};
    // Each thread walks one run of consecutive words; the runs are in
    // thread order.  Runs of `per` words (L4 / BLOCK) put the k-th word of
    // every run of a wave in one LDS bank (per = 32 at full windows: all 64
    // lanes of an in-order read on one bank).  Instead, the 32 threads of a
    // group share the group's 32*per words in runs of per + 1 words (the
    // last runs of a group are shorter, or empty), so at step k thread j of
    // the group reads bank (j*(per+1) + k) mod 32: a different bank for each
    // thread when per + 1 is odd.
    const uint32_t per = L4 / BLOCK;
    const uint32_t gbase = ((uint32_t)tid >> 5) * 32u * per, gend = gbase + 32u * per;
    const uint32_t i0 = min(gbase + ((uint32_t)tid & 31u) * (per + 1u), gend);
    const uint32_t iend = min(i0 + per + 1u, gend);
    // pre-pass: clip the run to [n_lo, n_hi] (only runs at the window's
    // ends need it) and add its words into bit-sliced counters: bit j
    // of v[k] is bit k of the number of words with bit j set (a run has
    // at most 33 words, so six planes hold every count)
    uint32_t v0 = 0, v1 = 0, v2 = 0, v3 = 0, v4 = 0, v5 = 0;
    auto add_word = [&](uint32_t w) {
        uint32_t c = w, t;
        t = v0 & c; v0 ^= c; c = t;
        t = v1 & c; v1 ^= c; c = t;
        t = v2 & c; v2 ^= c; c = t;
        t = v3 & c; v3 ^= c; c = t;
        t = v4 & c; v4 ^= c; c = t;
        v5 ^= c;
    };
    if ((int)(32u * i0) < bitlo || (int)(32u * iend) - 1 > bithi) {
        for (uint32_t i = i0; i < iend; i++) {
            const uint32_t w0 = bm[i];
            uint32_t w = w0;
            const int lo = (int)(32u * i);
            if (lo < bitlo) w &= (bitlo - lo >= 32) ? 0u : (0xFFFFFFFFu << (bitlo - lo));
            if (lo + 31 > bithi)
                w &= (bithi < lo) ? 0u
                     : (bithi - lo >= 31 ? 0xFFFFFFFFu : ((2u << (bithi - lo)) - 1u));
            if (w != w0) bm[i] = w;
            add_word(w);
        }
    } else {
        for (uint32_t i = i0; i < iend; i++) add_word(bm[i]);
    }
    // counts from the planes: all primes, p % 4 == 1 (bytes 0 and 2 of a
    // word have even k: residues {1,13,17,29} there, {7,11,19,23} at odd
    // k), and per wheel residue
    uint32_t cnt = 0, c1 = 0, m30[8];
#pragma unroll
    for (int o = 0; o < 8; o++) m30[o] = 0;
    {
        const uint32_t vp[6] = {v0, v1, v2, v3, v4, v5};
#pragma unroll
        for (int k = 0; k < 6; k++) {
            cnt += (uint32_t)__popc(vp[k]) << k;
            c1 += (uint32_t)__popc(vp[k] & 0x66996699u) << k;
#pragma unroll
            for (int o = 0; o < 8; o++)
                m30[o] += (uint32_t)__popc(vp[k] & (0x01010101u << o)) << k;
        }
    }
    // EMIT: exclusive prefix of the per-thread counts over the block
    // gives each thread its first output slot (ordered compaction)
    unsigned long long emit_at = 0;
    if constexpr (EMIT) {
        uint32_t incl = cnt;
#pragma unroll
        for (int s2 = 1; s2 < 64; s2 <<= 1) {
            const uint32_t y = __shfl_up(incl, s2);
            if (lane >= s2) incl += y;
        }
        if (lane == 63) S.tfirst[tid >> 6] = incl;
        __syncthreads();
        uint32_t woff = 0;
        for (int k = 0; k < (tid >> 6); k++) woff += S.tfirst[k];
        emit_at = io.emit_at + woff + incl - cnt;
        __syncthreads();
    }
    // walk: the primes of the run in increasing order
    uint32_t first = 0, prev = 0;
    {
        uint32_t i = i0;
        while (i < iend && bm[i] == 0u) i++;
        uint32_t prevoff = 0, rowb = 0;  // rowb = 4 * wheel bit of p_before
        if (i < iend) {
            const uint32_t b = __builtin_ctz(bm[i]);
            prevoff = 120u * i + 30u * (b >> 3) +
                      __builtin_amdgcn_perm(0x1D171311u, 0x0D0B0701u, (b & 7u) | 0x0C0C0C00u);
            first = prevoff + 1u;
            prev = cnt ? 1u : 0u;
        }
        uint32_t w = i < iend ? bm[i] : 0u, base = 120u * i;
        if constexpr (EMIT) {
            // every prime of the run in order, one iteration per prime
            // (cnt counts the run's set bits, so the refill never passes
            // iend)
            for (uint32_t k = 0; k < cnt; k++) {
                while (w == 0u) {
                    i++;
                    w = bm[i];
                    base += 120u;
                }
                const uint32_t b = __builtin_ctz(w);
                w &= w - 1u;
                // 30*(b>>3) + W30[b&7]; v_perm picks byte (b&7) of W30
                // and selector bytes 0x0C give zeros above it
                prevoff = base + 30u * (b >> 3) +
                          __builtin_amdgcn_perm(0x1D171311u, 0x0D0B0701u, (b & 7u) | 0x0C0C0C00u);
                a.emit_out[emit_at++] = base_num + prevoff;
            }
        } else {
            // one pass over the run's words from the first nonzero one,
            // the four wheel bytes (30 numbers each) of a word unrolled:
            // a nonzero byte's first prime pairs with the previous
            // byte's last one here (one LDS atomic: the q=30 row of
            // p_before for gaps below 100, the main histogram above;
            // the first prime of the run pairs with itself, gap 0, into
            // the dropped bin 0); the pairs inside the byte depend on the
            // byte value alone and are counted by value (S.bytes, only
            // for bytes with two primes or more), expanded at the flush.
            // A lane runs one iteration per word (not per byte), so
            // lanes with more primes do not hold the wave longer.  The
            // next word is loaded one iteration ahead.  Large-gap
            // candidates (g >= g_maybe, rare) branch to pair_large.
            char* const mqb = (lane & 1) ? (char*)S.mq2 : (char*)S.mq;
            char* const hb = (char*)S.hist;
            uint32_t wn = w;
            for (; i < iend; i++) {
                const uint32_t cw = wn;
                wn = bm[min(i + 1u, iend - 1u)];
                const uint32_t wb = 120u * i;
#pragma unroll
                for (uint32_t j = 0; j < 4; j++) {
                    const uint32_t v = (cw >> (8u * j)) & 0xFFu;
                    if (v) {
                        const uint32_t e = S.blut[v];
                        const uint32_t bb = wb + 30u * j;
                        const uint32_t g = bb + (e & 31u) - prevoff;  // even
                        char* const at = g < 100u ? mqb + (16u * g + rowb)
                                                  : hb + 2u * min(g, 2u * kHistBins - 2u);
                        atomicAdd(reinterpret_cast<uint32_t*>(at), 1u);
                        if (g >= pc.g_maybe) pair_large(pc, prevoff, g);
                        if (e & 0x8000u) atomicAdd(&S.bytes[v], 1u);
                        rowb = (e >> 10) & 0x1Cu;
                        prevoff = bb + ((e >> 5) & 31u);
                    }
                }
            }
        }
        prev = prev ? prevoff + 1u : 0u;
    }
    S.tfirst[tid] = first;
    S.tlast[tid] = prev;

    // per-window counts: wave reduction, then one LDS atomic per wave
    {
        uint32_t v[10] = {cnt, c1, m30[0], m30[1], m30[2], m30[3],
                          m30[4], m30[5], m30[6], m30[7]};
#pragma unroll
        for (int k = 0; k < 10; k++) {
            const uint32_t x = wave_sum63(v[k]);
            if (lane == 63 && x) atomicAdd(&S.cnt[k], (unsigned long long)x);
        }
    }
    __syncthreads();

    // pairs that cross thread runs
    if (first) {
        for (int s = tid - 1; s >= 0; s--) {
            const uint32_t l = S.tlast[s];
            if (l) {
                const uint32_t pb = l - 1u;
                const uint32_t g = first - 1u - pb;
                if (g < 100u) atomicAdd(&S.mq[((g >> 1) << 3) + (uint32_t)c_residx[pb % 30u]], 1u);
                else pair_wide(pc, pb, g);
                break;
            }
        }
    }
    if (tid == 0) {
        SegRec rc;
        rc.count = S.cnt[0];
        rc.first = 0;
        rc.last = 0;
        if (rc.count) {
            for (int s = 0; s < BLOCK; s++)
                if (S.tfirst[s]) { rc.first = base_num + S.tfirst[s] - 1u; break; }
            for (int s = BLOCK - 1; s >= 0; s--)
                if (S.tlast[s]) { rc.last = base_num + S.tlast[s] - 1u; break; }
            JobAcc* acc = io.acc;
            atomicAdd(&acc->count, S.cnt[0]);
            atomicAdd(&acc->c1, S.cnt[1]);
            for (int o = 0; o < 8; o++) atomicAdd(&acc->mod30[o], S.cnt[2 + o]);
            atomicMin(&acc->first, rc.first);
            atomicMax(&acc->last, rc.last);
        }
        *io.rec = rc;
    }
    __syncthreads();
}

// Large-prime buckets.  One workgroup per task: a tile of <= 256
// consecutive windows of one job and the primes >= bucket_limit of the
// task.  An owner task has every prime of its tile, so the tile's buckets
// have a single writer and the slots are handed out by LDS atomics alone;
// a shared task (a prime chunk of a tile too big for one workgroup) counts
// its hits per window in LDS first and reserves them with one global atomic
// per window on the bucket's cursor (the cursors start at 0).  Tasks are
// taken from a device counter, largest first (the host orders them).  Two ways to enumerate the
// composite multiples m*p (m >= 2, coprime to 30) of a prime in the tile:
//   class mode  thread per prime: one Barrett reduction, then the eight
//               residue classes in turn (first multiple of the class, then
//               stride p bytes).  Best when p is below the tile length
//               (several hits per class);
//   walk mode   thread per prime: one remainder (start_mod, or mod_u64)
//               gives the first multiple at or after the tile start, then
//               the thread steps
//               m -> m+1 (p/30 bytes and p%30 in the residue) and keeps the
//               multiples coprime to 30.  Best for p beyond the tile length,
//               where a class mode thread mostly finds nothing.
// COUNT writes the hits per window to bucket_cnt.  FILL starts each
// window's LDS cursor at its bucket's first slot (relative to the task's
// first bucket) and writes the entries
// (offset << 3) | bit; a bucket holds bucket_off[b+1] - bucket_off[b]
// entries: hits beyond that are not written, the window's count (written to
// bucket_cnt) still includes them and *overflow is set (the caller then
// re-runs with exact counts), so no write ever leaves the bucket.

// Walk mode: calls f(byte offset in the tile, wheel bit) for every hit of
// prime p in the tile of span bytes starting at number X; r = X mod p.
template <typename F>
__device__ __forceinline__ void walk_hits(uint64_t X, uint32_t p, uint32_t r, uint32_t span, F&& f) {
    // first multiple m*p >= X with m >= 2, minus X
    const uint32_t x = X < 2ull * p ? (uint32_t)(2ull * p - X) : (r ? p - r : 0u);
    // the tile holds 30*span < 2^30 numbers; most walk-mode primes of a
    // short tile have no multiple in it, and one beyond 30*span at most one
    const uint32_t L = 30u * span;
    if (x >= L) return;
    uint32_t b = x / 30u, rr = x - 30u * b;
    if (p >= L) {
        if ((kCoprime30 >> rr) & 1u) f(b, wheel_bit_of_residue(rr));
        return;
    }
    const uint32_t pd = p / 30u, pm = p - 30u * pd;
    while (b < span) {
        if ((kCoprime30 >> rr) & 1u) f(b, wheel_bit_of_residue(rr));
        b += pd;
        rr += pm;
        if (rr >= 30u) {
            rr -= 30u;
            b++;
        }
    }
}

// Class mode, the same for every hit.
template <typename F>
__device__ __forceinline__ void class_hits(uint64_t w0, const PrimeRec& pr, uint32_t span, F&& f) {
    const uint32_t p = pr.p;
    const uint32_t r = mod_u64(w0, p, pr.magic);
#pragma unroll 1
    for (int j = 0; j < 8; j++) {
        const uint32_t bit = (pr.perm >> (3 * j)) & 7u;
        for (uint32_t b = first_pos(p, r, j, w0); b < span; b += p) f(b, bit);
    }
}

// FILL in class mode, four hits per step: the four slot reservations (LDS
// atomics with return) and the four room reads are issued before any result
// is used, so their latencies overlap instead of adding up per hit.
__device__ __forceinline__ void bucket_fill_class(uint64_t w0, const PrimeRec& pr, uint32_t span,
                                                  uint32_t* cnt, const uint32_t* lim,
                                                  uint32_t* entries, int log2w, uint32_t wmask) {
    const uint32_t p = pr.p;
    const uint32_t r = mod_u64(w0, p, pr.magic);
#pragma unroll 1
    for (int j = 0; j < 8; j++) {
        const uint32_t bit = (pr.perm >> (3 * j)) & 7u;
        uint32_t b = first_pos(p, r, j, w0);
        while ((uint64_t)b + 3ull * p < span) {
            uint32_t bb[4], sl[4], lm[4];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                bb[q] = b + (uint32_t)q * p;
                sl[q] = atomicAdd(&cnt[bb[q] >> log2w], 1u);
            }
#pragma unroll
            for (int q = 0; q < 4; q++) lm[q] = lim[bb[q] >> log2w];
#pragma unroll
            for (int q = 0; q < 4; q++)
                if (sl[q] < lm[q]) entries[sl[q]] = ((bb[q] & wmask) << 3) | bit;
            b += 4u * p;
        }
        for (; b < span; b += p) {
            const uint32_t w = b >> log2w;
            const uint32_t slot = atomicAdd(&cnt[w], 1u);
            if (slot < lim[w]) entries[slot] = ((b & wmask) << 3) | bit;
        }
    }
}

// g(record) for every prime of [i0, i1), NPF per thread at a time: the NPF
// prime records are loaded before any is used, so one load latency covers
// NPF primes.
template <int NPF, typename G>
__device__ __forceinline__ void bucket_sweep_primes(uint32_t i0, uint32_t i1,
                                                    const PrimeRec* __restrict__ primes, G&& g) {
    if (i1 <= i0) return;
    const uint32_t n = i1 - i0;
    const uint32_t last = i1 - 1u;
    for (uint32_t k0 = threadIdx.x; k0 < n; k0 += (uint32_t)NPF * blockDim.x) {
        PrimeRec pr[NPF];
#pragma unroll
        for (int q = 0; q < NPF; q++) {
            const uint32_t i = i0 + k0 + q * blockDim.x;
            pr[q] = primes[i < last ? i : last];
        }
#pragma unroll
        for (int q = 0; q < NPF; q++)
            if (k0 + q * blockDim.x < n) g(pr[q]);
    }
}

// g(p) for every prime of [i0, i1) from the packed 32-bit table, NPF per
// thread at a time.
template <int NPF, typename G>
__device__ __forceinline__ void bucket_sweep_p32(uint32_t i0, uint32_t i1,
                                                 const uint32_t* __restrict__ p32, G&& g) {
    if (i1 <= i0) return;
    const uint32_t n = i1 - i0;
    const uint32_t last = i1 - 1u;
    for (uint32_t k0 = threadIdx.x; k0 < n; k0 += (uint32_t)NPF * blockDim.x) {
        uint32_t pv[NPF];
#pragma unroll
        for (int q = 0; q < NPF; q++) {
            const uint32_t i = i0 + k0 + q * blockDim.x;
            pv[q] = p32[i < last ? i : last];
        }
#pragma unroll
        for (int q = 0; q < NPF; q++)
            if (k0 + q * blockDim.x < n) g(pv[q]);
    }
}

// Walk mode over the primes [i0, i1) of a tile (t: its first number): from
// the 4-byte packed table with start_mod where it applies (a third of the
// record bytes per prime and no 64-bit multiplies), else from the records
// with mod_u64.
template <int NPF, typename F>
__device__ __forceinline__ void walk_sweep(const ModStart& t, uint32_t i0, uint32_t i1,
                                           const PrimeRec* __restrict__ primes,
                                           const uint32_t* __restrict__ p32, uint32_t span, F&& f) {
    if (t.fp) {
        bucket_sweep_p32<2 * NPF>(i0, i1, p32, [&](uint32_t p) {
            walk_hits(t.X, p, start_mod(t, p), span, f);
        });
    } else {
        bucket_sweep_primes<NPF>(i0, i1, primes, [&](const PrimeRec& pr) {
            walk_hits(t.X, pr.p, mod_u64(t.X, pr.p, pr.magic), span, f);
        });
    }
}

// One bucket task run by the whole
// workgroup; cnt and lim are LDS arrays of kBucketTileMax entries.  Every
// thread calls it; it ends with a barrier, so cnt and lim may be reused.
// HOST_OFFSETS: bucket_off is host memory that the host writes while the
// kernel runs, read with system-scope loads (they bypass the GPU caches).
// NPF: prime records in flight per thread (more where the workgroup's
// occupancy is low and the record loads' latency shows).
template <bool FILL, bool HOST_OFFSETS = false, int NPF = 4>
__device__ __forceinline__ void bucket_task(const BucketTask& T, const PrimeRec* __restrict__ primes,
                                            const uint32_t* __restrict__ p32,
                                            uint32_t* __restrict__ bucket_cnt,
                                            const unsigned long long* __restrict__ bucket_off_p,
                                            uint32_t* __restrict__ entries,
                                            uint32_t* __restrict__ overflow, int log2w,
                                            uint32_t* cnt, uint32_t* lim) {
    struct Offsets {
        const unsigned long long* p;
        __device__ unsigned long long operator[](uint32_t i) const {
            if constexpr (HOST_OFFSETS)
                return __hip_atomic_load(const_cast<unsigned long long*>(p + i), __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_SYSTEM);
            else
                return p[i];
        }
    };
    const Offsets bucket_off{bucket_off_p};
    const uint32_t tid = threadIdx.x;
    const uint32_t wmask = (1u << log2w) - 1u;
    const uint32_t span = T.nwin << log2w;  // bytes; at most 2^25
    auto count = [&](uint32_t b, uint32_t) { atomicAdd(&cnt[b >> log2w], 1u); };
    const ModStart ts = mod_start(30ull * T.w0, T.i1 > T.iw ? p32[T.iw] : 1u,
                                  T.i1 > T.iw ? p32[T.i1 - 1u] : 0u);
    // slots count from the task's first bucket, so the LDS cursors stay
    // 32-bit however many entries the round holds
    const unsigned long long ebase = FILL ? bucket_off[T.b0] : 0ull;
    uint32_t* const ent = FILL ? entries + ebase : nullptr;
    for (uint32_t i = tid; i < T.nwin; i += blockDim.x) {
        if (FILL && !T.shared) {
            cnt[i] = (uint32_t)(bucket_off[T.b0 + i] - ebase);
            lim[i] = (uint32_t)(bucket_off[T.b0 + i + 1] - ebase);
        } else {
            cnt[i] = 0;
        }
    }
    __syncthreads();
    if (FILL && T.shared) {
        // count, then reserve this task's slots in each window's bucket
        bucket_sweep_primes<NPF>(T.i0, T.iw, primes, [&](const PrimeRec& pr) {
            class_hits(T.w0, pr, span, count);
        });
        walk_sweep<NPF>(ts, T.iw, T.i1, primes, p32, span, count);
        __syncthreads();
        for (uint32_t i = tid; i < T.nwin; i += blockDim.x) {
            const unsigned long long o0 = bucket_off[T.b0 + i], o1 = bucket_off[T.b0 + i + 1];
            const uint32_t start = cnt[i] ? atomicAdd(&bucket_cnt[T.b0 + i], cnt[i]) : 0u;
            if ((unsigned long long)start + cnt[i] > o1 - o0) atomicOr(overflow, 1u);
            cnt[i] = (uint32_t)(o0 - ebase) + start;
            lim[i] = (uint32_t)(o1 - ebase);
        }
        __syncthreads();
    }
    if (!FILL) {
        bucket_sweep_primes<NPF>(T.i0, T.iw, primes, [&](const PrimeRec& pr) {
            class_hits(T.w0, pr, span, count);
        });
        walk_sweep<NPF>(ts, T.iw, T.i1, primes, p32, span, count);
    } else {
        bucket_sweep_primes<NPF>(T.i0, T.iw, primes, [&](const PrimeRec& pr) {
            bucket_fill_class(T.w0, pr, span, cnt, lim, ent, log2w, wmask);
        });
        walk_sweep<NPF>(ts, T.iw, T.i1, primes, p32, span, [&](uint32_t b, uint32_t bit) {
            const uint32_t w = b >> log2w;
            const uint32_t slot = atomicAdd(&cnt[w], 1u);
            if (slot < lim[w]) ent[slot] = ((b & wmask) << 3) | bit;
        });
    }
    __syncthreads();
    for (uint32_t i = tid; i < T.nwin; i += blockDim.x) {
        if (T.shared) {
            if (!FILL && cnt[i]) atomicAdd(&bucket_cnt[T.b0 + i], cnt[i]);
        } else if (FILL) {
            bucket_cnt[T.b0 + i] = cnt[i] - (uint32_t)(bucket_off[T.b0 + i] - ebase);
            if (cnt[i] > lim[i]) atomicOr(overflow, 1u);
        } else {
            bucket_cnt[T.b0 + i] = cnt[i];
        }
    }
    __syncthreads();  // cnt and lim are reused by the next task
}

}  // namespace voss
