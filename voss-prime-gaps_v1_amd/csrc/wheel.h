// wheel.h -- wheel-30 layout and the device-side records shared by the
// sieve kernels and the host planner.
//
// Bit layout (same meaning as the reference, src/voss_v7_live.cu:37-38,
// :189): wheel byte k holds the 8 numbers 30k + W30[o], o = 0..7, bit o of
// the byte.  Bytes are packed little-endian into 32-bit LDS words, so bit
// (8*j + o) of word i is the number 30*(4i + j) + W30[o].
#pragma once
#include <stdint.h>

namespace voss {

constexpr int kW30[8] = {1, 7, 11, 13, 17, 19, 23, 29};

// Device histogram geometry.  Every prime gap below 2^64 is < 1600
// (maximal gaps are known to 2^64), so even gaps g < 2048 cover all of
// them; a larger gap raises VOSS_ERANGE instead of being dropped.
constexpr int kHistBins = 1024;       // bin g/2, g < 2048
constexpr int kModqBins = 50;         // bin g/2, g < 100
constexpr int kModqRows = 10;         // q=6 {1,5}, q=30 {1,7,...,29}
// In LDS only the 8 q=30 rows are kept: the q=6 rows and the main
// histogram below g=100 are sums of them for pairs inside a window.
constexpr int kLargeMin = 101;        // merit only for g > 100 (:290)
constexpr int kLargeAlways = 500;     // LARGE_GAP_THRESHOLD (:33)

// Presieve: primes 7..37 are removed by ANDing period patterns held in
// LDS: one table of prod(p) words per group (the word pattern of a prime
// p repeats every p words).  4 tables, 3142 words = 12.6 KB.  Primes 41..73
// had one table each; marked as tier A instead (a wave per prime) they cost
// about what their table reads did and free 2 KB of LDS (1e11: presieve 13.1
// -> 4.8 us per window, marking 67.2 -> 75.0 us).
constexpr int kPresieveGroups = 4;
constexpr int kPresievePrimes = 9;    // 7,11,...,37
constexpr int kPresieveGroupPrimes[kPresieveGroups][3] = {
    {7, 11, 13}, {17, 19, 1}, {23, 29, 1}, {31, 37, 1}};
constexpr int kPresieveWords = 1001 + 323 + 667 + 1147;
// Each table is followed by a copy of its first word, so two consecutive
// pattern words can be read without a wrap check.
constexpr int kPresieveWordsPadded = kPresieveWords + kPresieveGroups;

// One unit of sieve work: an LDS window of `len` wheel bytes starting at
// byte w0 (w0 % 4 == 0, len % 4096 == 0), of which the numbers in
// [n_lo, n_hi] are counted.  Pairs of primes that lie in two different
// windows are joined by the stitch kernel; they count as "in-segment"
// (full rule) only when both windows carry the same lseg.
struct SegDesc {
    uint64_t w0;
    uint64_t n_lo;
    uint64_t n_hi;
    uint32_t len;
    uint32_t nprimes;  // base primes marked per window: p*p <= last number
                       // of the window, and p below the bucket limit when
                       // the window has a bucket
    uint32_t job;      // output slot (interval index in batch mode)
    uint32_t lseg;     // logical segment id (plain mode)
    uint32_t bucket;   // bucket of this window's large-prime hits, or kNoBucket
    uint32_t ibig;     // index of the first base prime >= len (at most one
                       // hit per residue class in the window from there on)
};
constexpr uint32_t kNoBucket = 0xFFFFFFFFu;

// Large-prime bucket fill: one task = the primes [i0, i1) of one job over
// `nwin` consecutive grid windows starting at byte w0 (buckets b0..).
// A bucket entry is (byte offset in the window << 3) | wheel bit.
// One bucket task: a tile of nwin consecutive windows of one job (buckets
// b0 .. b0+nwin-1) and large primes [i0, iw) in class mode, [iw, i1) in
// walk mode.  An owner task holds all the tile's primes, so its buckets see
// no other writer; a tile with too much work for one workgroup is split
// into shared tasks (prime chunks) that reserve their slots with global
// atomics on the buckets' cursors.
struct BucketTask {
    uint64_t w0;
    uint32_t b0;
    uint32_t nwin;
    uint32_t i0;
    uint32_t iw;
    uint32_t i1;
    uint32_t shared;  // 1: one of several tasks of its tile
};
constexpr int kBucketTileMax = 256;  // windows per task (LDS counters)

// Per base prime, one 16-byte record read with a single load: the prime,
// in perm the wheel bit of p*W30[j] for each residue class j (bits 3j..3j+2)
// and p % 30 (bits 24..28), and the Barrett constant floor(2^64 / p).
struct PrimeRec {
    uint32_t p;
    uint32_t perm;
    unsigned long long magic;
};

// Written by the sieve for every descriptor; read by the stitch kernel.
struct SegRec {
    unsigned long long count;
    unsigned long long first;  // 0 if count == 0
    unsigned long long last;
};

// Per-job accumulators in HBM (all unsigned 64-bit, summed with atomics,
// so the result is independent of scheduling order).
struct JobAcc {
    unsigned long long hist[kHistBins];             // all pairs, bin g/2
    unsigned long long modq[kModqRows * kModqBins]; // in-segment, g < 100
    unsigned long long count;                       // wheel primes
    unsigned long long c1;                          // p % 4 == 1
    unsigned long long mod30[8];                    // by wheel residue
    unsigned long long first;                       // min wheel prime
    unsigned long long last;                        // max wheel prime
    unsigned long long large_count;                 // large gaps emitted
    unsigned long long error;                       // bit 0: gap >= 2048
};

// Large-gap candidate emitted by the device; the host turns it into the
// reference's LargeGap {p_after, gap, merit}.
struct LargeRaw {
    unsigned long long p_before;
    uint32_t gap;
    uint32_t job;
};

}  // namespace voss
