// voss_v7_live -- the sieve+gap program of the drop-in, over libvoss_amd.
//
//   voss_v7_live N SEG_NUM [--device D] [--max-large K]
//
// Process boundary of the reference's src/voss_v7_live.cu main() (run with
// cwd = results/ by src/voss_master.py): the same stdout lines that its
// driver parses ("Total primes: <pi>", src/voss_master.py:170), the same
// output files in cwd (src/voss_v7_live.cu:707-827) and the same exit
// codes (0 on success or stop-file, EXIT_FAILURE on a device error).  N and
// SEG_NUM arrive as arguments instead of being patched into the source.
//
// Checkpoint/resume (checkpoint.bin, stop.txt) keeps the reference's file
// layout and cadence (:42-89, :599-621): logical segments are run in
// groups that end at the reference's checkpoint segments.
#include <cerrno>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "voss/voss.h"

namespace {

constexpr int64_t kMaxGap = VOSS_MAX_GAP;
constexpr int64_t kMaxLarge = 50000000;  // MAX_LARGE_GAPS (:34)

// checkpoint.bin: the reference's CheckpointState (src/voss_v7_live.cu:42-54,
// 800,064 bytes, native layout) followed by an optional extension with the
// state the reference does not keep (mod-q counts, per-residue gap
// histogram, large gaps).  The reference reads the first part only.
constexpr int64_t kCkptMagic = 0x564F53534F4D4ELL;
constexpr uint64_t kExtMagic = 0x31444D4153534F56ULL;  // "VOSSAMD1"

struct CheckpointState {
    int64_t magic;
    int64_t N;
    int64_t SEG_NUM;
    int64_t next_segment;
    uint64_t last_prime;
    uint64_t total_primes;
    unsigned long long class1_total;
    unsigned long long class3_total;
    int64_t gaps_total[VOSS_MAX_GAP];
};
static_assert(sizeof(CheckpointState) == 800064, "reference checkpoint layout");

struct RunState {
    CheckpointState ck;
    bool has_ext = false;
    uint64_t modq[VOSS_Q_SLOTS] = {};
    int64_t modq_gap[VOSS_MODQ_ROWS * VOSS_MODQ_MAX_GAP] = {};
    uint64_t large_count = 0;
    std::vector<voss_large_gap> large;
};

bool load_checkpoint(const char* path, int64_t N, int64_t SEG, RunState* st) {
    FILE* f = std::fopen(path, "rb");
    if (!f) return false;
    bool ok = std::fread(&st->ck, sizeof(CheckpointState), 1, f) == 1 &&
              st->ck.magic == kCkptMagic && st->ck.N == N && st->ck.SEG_NUM == SEG;  // :56-68
    if (ok) {
        uint64_t m = 0, nl = 0;
        if (std::fread(&m, 8, 1, f) == 1 && m == kExtMagic &&
            std::fread(st->modq, sizeof(st->modq), 1, f) == 1 &&
            std::fread(st->modq_gap, sizeof(st->modq_gap), 1, f) == 1 &&
            std::fread(&st->large_count, 8, 1, f) == 1 && std::fread(&nl, 8, 1, f) == 1) {
            st->large.resize(nl);
            if (nl == 0 || std::fread(st->large.data(), sizeof(voss_large_gap), nl, f) == nl)
                st->has_ext = true;
        }
        if (!st->has_ext) {
            std::memset(st->modq, 0, sizeof(st->modq));
            std::memset(st->modq_gap, 0, sizeof(st->modq_gap));
            st->large_count = 0;
            st->large.clear();
        }
    }
    std::fclose(f);
    return ok;
}

void save_checkpoint(const char* path, const RunState& st) {
    FILE* f = std::fopen(path, "wb");
    if (!f) return;  // the reference ignores a failed checkpoint write (:85-86)
    const uint64_t nl = st.large.size();
    std::fwrite(&st.ck, sizeof(CheckpointState), 1, f);
    std::fwrite(&kExtMagic, 8, 1, f);
    std::fwrite(st.modq, sizeof(st.modq), 1, f);
    std::fwrite(st.modq_gap, sizeof(st.modq_gap), 1, f);
    std::fwrite(&st.large_count, 8, 1, f);
    std::fwrite(&nl, 8, 1, f);
    if (nl) std::fwrite(st.large.data(), sizeof(voss_large_gap), nl, f);
    std::fclose(f);
}

[[noreturn]] void die(const char* what) {
    std::fprintf(stderr, "VOSS error in %s: %s\n", what, voss_last_error());
    std::exit(EXIT_FAILURE);
}

bool parse_u64(const char* s, uint64_t* v) {
    if (!s || !*s) return false;
    char* end = nullptr;
    errno = 0;
    unsigned long long x = std::strtoull(s, &end, 10);
    if (errno || *end) {
        // accept "1e11"-style input
        double d = std::strtod(s, &end);
        if (*end || d < 0 || d > 1.8e19) return false;
        x = (unsigned long long)std::llround(d);
    }
    *v = x;
    return true;
}

// A duration as the progress line shows it: seconds below a minute, then
// minutes below an hour, then hours, one decimal.
std::string duration_text(double seconds) {
    static const struct { double limit, unit; char tag; } scales[] = {
        {60.0, 1.0, 's'}, {3600.0, 60.0, 'm'}, {1e300, 3600.0, 'h'}};
    char out[32];
    for (const auto& sc : scales)
        if (seconds < sc.limit) {
            std::snprintf(out, sizeof out, "%5.1f %c", seconds / sc.unit, sc.tag);
            break;
        }
    return out;
}

// The "\r[ %] seg i/n | primes | last p | elapsed | ETA" line the reference
// program rewrites after every segment group (stdout contract,
// src/voss_v7_live.cu:317-335); the ETA assumes the rate so far.
void show_progress(int64_t done, int64_t of, uint64_t primes, uint64_t last, double elapsed) {
    const double per_s = elapsed > 0.1 ? (double)done / elapsed : 0.0;
    const double left = per_s > 0.0 ? (double)(of - done) / per_s : 0.0;
    std::fprintf(stdout,
                 "\r[%3d%%] seg %lld/%lld | primes: %-13llu | last p: %-16llu | elapsed: %s | ETA: %s   ",
                 (int)(done * 100 / of), (long long)done, (long long)of, (unsigned long long)primes,
                 (unsigned long long)last, duration_text(elapsed).c_str(), duration_text(left).c_str());
    std::fflush(stdout);
}

// Moments of the gap distribution printed after a run: number of gaps,
// sum of g N(g), mean, standard deviation, skewness and excess kurtosis of
// g weighted by N(g), in double precision, terms added in increasing g (the
// printed digits follow from that order of operations).
struct GapMoments {
    int64_t count = 0, weighted = 0;
    double mean = 0, sd = 0, skew = 0, kurt = 0;
};

GapMoments gap_moments(const int64_t* hist, int bins) {
    GapMoments m;
    for (int g = 1; g < bins; g++) {
        m.count += hist[g];
        m.weighted += (int64_t)g * hist[g];
    }
    m.mean = (double)m.weighted / (double)m.count;
    double c2 = 0, c3 = 0, c4 = 0;  // central sums: N(g) (g - mean)^k
    for (int g = 1; g < bins; g++) {
        if (!hist[g]) continue;
        const double dev = (double)g - m.mean;
        const double t = (double)hist[g] * dev;
        c2 += t * dev;
        c3 += t * dev * dev;
        c4 += t * dev * dev * dev;
    }
    const double var = c2 / (double)m.count;
    m.sd = std::sqrt(var);
    m.skew = c3 / (double)m.count / (var * m.sd);
    m.kurt = c4 / (double)m.count / (var * var) - 3.0;
    return m;
}

}  // namespace

int main(int argc, char** argv) {
    setvbuf(stdout, nullptr, _IONBF, 0);
    uint64_t N = 0, SEG_NUM = 30000000000ull, max_large = kMaxLarge;
    int device = 0;
    std::vector<const char*> pos;
    for (int i = 1; i < argc; i++) {
        std::string a = argv[i];
        if (a == "--device" && i + 1 < argc) device = std::atoi(argv[++i]);
        else if (a == "--max-large" && i + 1 < argc) {
            if (!parse_u64(argv[++i], &max_large)) return 2;
        } else pos.push_back(argv[i]);
    }
    if (pos.empty() || !parse_u64(pos[0], &N) || (pos.size() > 1 && !parse_u64(pos[1], &SEG_NUM))) {
        std::fprintf(stderr, "usage: voss_v7_live N [SEG_NUM] [--device D] [--max-large K]\n");
        return 2;
    }
    if (N < 5 || SEG_NUM == 0 || SEG_NUM % 30 != 0) {
        std::fprintf(stderr, "N must be >= 5 and SEG_NUM a positive multiple of 30\n");
        return 2;
    }
    const int64_t NUM_SEG = (int64_t)((N + SEG_NUM - 1) / SEG_NUM);

    std::printf("=============================================================\n");
    std::printf("  VOSS-Wheel30 v7 LIVE  (MI355X / gfx950, LDS-window sieve)\n");
    std::printf("=============================================================\n");
    std::printf("N         = %lld\n", (long long)N);
    std::printf("SEG_NUM   = %lld\n", (long long)SEG_NUM);
    std::printf("NUM_SEG   = %lld\n", (long long)NUM_SEG);
    std::printf("=============================================================\n\n");

    voss_ctx* ctx = nullptr;
    if (voss_create(device, &ctx) != VOSS_OK) die("voss_create");
    char name[256];
    int cus = 0;
    double gb = 0;
    voss_device_name(ctx, name, sizeof(name), &cus, &gb);
    std::printf("GPU: %s (%d CUs, %.1f GB)\n\n", name, cus, gb);
    const int limit = (int)std::sqrt((double)N) + 1;  // :357
    std::printf("Base primes (>=7) up to sqrt(N)=%d\n\n", limit);
    static RunState st;
    std::memset(&st.ck, 0, sizeof(st.ck));
    int64_t start_seg = 0;
    if (load_checkpoint("checkpoint.bin", (int64_t)N, (int64_t)SEG_NUM, &st)) {
        start_seg = st.ck.next_segment;
        std::printf(">>> RESUMING FROM SEGMENT %lld / %lld <<<\n", (long long)start_seg,
                    (long long)NUM_SEG);
    } else {
        std::memset(&st.ck, 0, sizeof(st.ck));
        st.ck.magic = kCkptMagic;
        st.ck.N = (int64_t)N;
        st.ck.SEG_NUM = (int64_t)SEG_NUM;
        st.ck.last_prime = 5;    // :404
        st.ck.total_primes = 3;
        std::printf(">>> STARTING FRESH <<<\n");
    }

    voss_result* r = (voss_result*)std::calloc(1, sizeof(voss_result));
    const uint64_t cap = std::min<uint64_t>(max_large, (uint64_t)kMaxLarge);
    std::vector<voss_large_gap> large(cap ? cap : 1);
    r->large = large.data();
    r->large_capacity = cap;

    // logical segments in the reference's checkpoint groups: segment 0,
    // then up to each (seg+1) % 5 == 0, then the last (:599-602)
    double total_s = 0, sieve_ms = 0, dev_ms = 0;
    for (int64_t seg = start_seg; seg < NUM_SEG;) {
        int64_t end = seg + 1;
        while (end < NUM_SEG && !(end - 1 == 0 || end % 5 == 0)) end++;
        if (voss_run_plain_range(ctx, N, SEG_NUM, (uint64_t)seg, (uint64_t)end, r) != VOSS_OK)
            die("voss_run_plain_range");
        total_s += r->seconds;
        sieve_ms += r->sieve_ms;
        dev_ms += r->device_ms;
        if (r->total_primes) {
            // the join of two groups crosses a SEG_NUM boundary: main
            // histogram only, like the reference's host boundary gap (:559-562)
            const uint64_t d0 = r->first_prime - st.ck.last_prime;
            if (d0 > 0 && d0 < (uint64_t)kMaxGap) st.ck.gaps_total[d0]++;
            st.ck.last_prime = r->last_prime;
            st.ck.total_primes += r->total_primes;
        }
        for (int g = 1; g < kMaxGap; g++) st.ck.gaps_total[g] += r->gaps[g];
        st.ck.class1_total += r->c1;
        st.ck.class3_total += r->c3;
        for (int i = 0; i < VOSS_Q_SLOTS; i++) st.modq[i] += r->modq[i];
        for (int i = 0; i < VOSS_MODQ_ROWS * VOSS_MODQ_MAX_GAP; i++) st.modq_gap[i] += r->modq_gap[i];
        st.large_count += r->large_count;
        const uint64_t nkeep = std::min<uint64_t>(r->large_count, cap);
        for (uint64_t i = 0; i < nkeep && st.large.size() < cap; i++) st.large.push_back(large[i]);
        st.ck.next_segment = end;
        save_checkpoint("checkpoint.bin", st);
        FILE* stop = std::fopen("stop.txt", "r");
        if (stop) {  // :603-620
            std::fclose(stop);
            std::remove("stop.txt");
            std::printf("\n>>> STOP-FILE detected at seg %lld — exiting cleanly <<<\n",
                        (long long)(end - 1));
            voss_destroy(ctx);
            return 0;
        }
        show_progress(end, NUM_SEG, st.ck.total_primes, st.ck.last_prime, total_s);
        seg = end;
    }
    std::printf("\n\n");
    std::remove("checkpoint.bin");

    // results in the reference's arrays
    std::memset(r, 0, sizeof(*r));
    std::memcpy(r->gaps, st.ck.gaps_total, sizeof(r->gaps));
    r->gaps[1] += 1;  // (2,3)  :634
    r->gaps[2] += 1;  // (3,5)  :635
    r->total_primes = st.ck.total_primes;
    r->last_prime = st.ck.last_prime;
    r->c1 = st.ck.class1_total;
    r->c3 = st.ck.class3_total;
    std::memcpy(r->modq, st.modq, sizeof(r->modq));
    std::memcpy(r->modq_gap, st.modq_gap, sizeof(r->modq_gap));
    r->large_count = st.large_count;
    r->seconds = total_s;
    r->sieve_ms = sieve_ms;
    r->device_ms = dev_ms;
    // canonical order: merit desc, position asc
    std::sort(st.large.begin(), st.large.end(), [](const voss_large_gap& a, const voss_large_gap& b) {
        if (a.merit != b.merit) return a.merit > b.merit;
        return a.position < b.position;
    });
    large.assign(st.large.begin(), st.large.end());
    int64_t* gaps = r->gaps;
    std::printf("=============================================================\n");
    std::printf("TOTAL TIME: %.2f s (%.2f min)\n", total_s, total_s / 60.0);
    std::printf("Total primes: %llu\n", (unsigned long long)r->total_primes);
    std::printf("=============================================================\n\n");

    std::printf("=== Timing Breakdown ===\n");
    std::printf("Sieve+extract+gaps (fused kernel): %10.3f s\n", r->sieve_ms / 1000.0);
    std::printf("Stitch (window edges):             %10.3f s\n", (r->device_ms - r->sieve_ms) / 1000.0);
    std::printf("Host (plan, copies, finalise):     %10.3f s\n\n", r->seconds - r->device_ms / 1000.0);

    // summary block (stdout contract, src/voss_v7_live.cu:652-705)
    const GapMoments mo = gap_moments(gaps, kMaxGap);
    std::printf("=== Gap Statistics ===\n");
    std::printf("Total gaps:  %lld\n", (long long)mo.count);
    std::printf("Mean gap:    %.4f\n", mo.mean);
    std::printf("Std dev:     %.4f\n", mo.sd);
    std::printf("Skewness:    %.4f\n", mo.skew);
    std::printf("Kurtosis:    %.4f\n\n", mo.kurt);
    std::printf("=== Gap Counts ===\n");
    for (int g : {1, 2, 4, 6, 12, 30})
        std::printf("Gap %d:%s%lld\n", g, g < 10 ? "  " : " ", (long long)gaps[g]);
    std::printf("\n");

    // pi(x;4,1) and pi(x;4,3) count 5 and 3 besides the wheel primes
    const unsigned long long pi41 = r->c1 + 1, pi43 = r->c3 + 1;
    const long long bias = (long long)pi43 - (long long)pi41;
    std::printf("=== Chebyshev Bias ===\n");
    std::printf("pi(x;4,1) = %llu\n", pi41);
    std::printf("pi(x;4,3) = %llu\n", pi43);
    std::printf("Diff      = %lld\n\n", bias);
    const double p6p2 = (double)gaps[6] / (double)gaps[2];  // also the hl_trend row
    std::printf("=== Gap Repulsion ===\n");
    std::printf("P(6)/P(2) = %.6f\n\n", p6p2);

    // the three identities every complete run satisfies
    const uint64_t pi = r->total_primes;
    const bool ok1 = mo.count == (int64_t)(pi - 1);
    const bool ok2 = mo.weighted == (int64_t)(r->last_prime - 2);
    const bool ok3 = pi41 + pi43 + 1 == pi;
    std::printf("=== Self-Consistency Checks ===\n");
    std::printf("[Check 1] sum(N(g)) = pi(N) - 1\n");
    std::printf("  sum(N(g))  = %lld\n", (long long)mo.count);
    std::printf("  pi(N) - 1  = %lld\n", (long long)(pi - 1));
    std::printf("  -> %s\n", ok1 ? "PASS" : "FAIL");
    std::printf("[Check 2] sum(g*N(g)) = p_last - 2\n");
    std::printf("  sum(g*N(g)) = %lld\n", (long long)mo.weighted);
    std::printf("  p_last - 2  = %lld\n", (long long)(r->last_prime - 2));
    std::printf("  -> %s\n", ok2 ? "PASS" : "FAIL");
    std::printf("[Check 3] pi(4,1) + pi(4,3) + 1 = pi(N)\n");
    std::printf("  pi41+pi43+1 = %llu\n", pi41 + pi43 + 1);
    std::printf("  pi(N)       = %llu\n", (unsigned long long)pi);
    std::printf("  -> %s\n\n", ok3 ? "PASS" : "FAIL");

    // ---- files (formats of :707-827)
    FILE* fp = std::fopen("gap_histogram.csv", "w");
    if (fp) {
        std::fprintf(fp, "gap,count\n");
        for (int g = 1; g < kMaxGap; g++)
            if (gaps[g] > 0) std::fprintf(fp, "%d,%lld\n", g, (long long)gaps[g]);
        std::fclose(fp);
        std::printf("OK gap_histogram.csv\n");
    }
    fp = std::fopen("hl_trend.csv", "w");
    if (fp) {
        std::fprintf(fp, "N,P6_over_P2\n");
        static const long long HL_NS[] = {1000000000LL, 10000000000LL, 100000000000LL,
                                          1000000000000LL, 10000000000000LL};
        static const double HL_RS[] = {1.778300, 1.801800, 1.820800, 1.836600, 1.849700};
        int already = 0;
        for (int i = 0; i < 5; i++) {
            std::fprintf(fp, "%lld,%.6f\n", HL_NS[i], HL_RS[i]);
            if ((long long)N == HL_NS[i]) already = 1;
        }
        if (!already) std::fprintf(fp, "%lld,%.6f\n", (long long)N, p6p2);
        std::fclose(fp);
        std::printf("OK hl_trend.csv (hybrid)\n");
    }
    fp = std::fopen("modq_counts.csv", "w");
    if (fp) {
        static const int QV[2] = {6, 30}, QO[2] = {0, 6};
        std::fprintf(fp, "q,a,count\n");
        for (int qi = 0; qi < 2; qi++)
            for (int a = 0; a < QV[qi]; a++)
                std::fprintf(fp, "%d,%d,%llu\n", QV[qi], a, (unsigned long long)r->modq[QO[qi] + a]);
        std::fclose(fp);
        std::printf("OK modq_counts.csv\n");
    }
    fp = std::fopen("modq_gap_hist.csv", "w");
    if (fp) {
        static const int QV[2] = {6, 30}, QOFF[2] = {0, 2}, QN[2] = {2, 8};
        static const int QRES[2][8] = {{1, 5, 0, 0, 0, 0, 0, 0}, {1, 7, 11, 13, 17, 19, 23, 29}};
        std::fprintf(fp, "q,residue,gap,count\n");
        for (int qi = 0; qi < 2; qi++)
            for (int ri = 0; ri < QN[qi]; ri++)
                for (int g = 1; g < VOSS_MODQ_MAX_GAP; g++) {
                    int64_t c = r->modq_gap[(QOFF[qi] + ri) * VOSS_MODQ_MAX_GAP + g];
                    if (c > 0) std::fprintf(fp, "%d,%d,%d,%lld\n", QV[qi], QRES[qi][ri], g, (long long)c);
                }
        std::fclose(fp);
        std::printf("OK modq_gap_hist.csv\n");
    }
    fp = std::fopen("chebyshev.csv", "w");
    if (fp) {
        std::fprintf(fp, "N,pi_4_1,pi_4_3,difference\n");
        std::fprintf(fp, "%lld,%llu,%llu,%lld\n", (long long)N, pi41, pi43, bias);
        std::fclose(fp);
        std::printf("OK chebyshev.csv\n");
    }
    {
        std::vector<int32_t> gs(VOSS_GAP_SEQ_SIZE, 0);
        std::vector<uint64_t> ps(VOSS_GAP_SEQ_SIZE, 0);
        uint64_t rows = 0;
        const int rc = voss_sample_plain(ctx, N, SEG_NUM, gs.data(), ps.data(), &rows);
        if (rc != VOSS_OK) die("voss_sample_plain");
        // after a resume the reference holds samples of this process's
        // segments only (gap_seq_d is filled per processed segment, :526-547)
        const int64_t per_seg = VOSS_GAP_SEQ_SIZE / NUM_SEG;
        for (int64_t i = 0; i < start_seg * per_seg && i < VOSS_GAP_SEQ_SIZE; i++) {
            gs[i] = 0;
            ps[i] = 0;
        }
        fp = std::fopen("gap_sequence.csv", "w");
        if (fp) {
            std::fprintf(fp, "index,gap,p\n");
            if (rc == VOSS_OK)
                for (int i = 0; i < VOSS_GAP_SEQ_SIZE; i++)
                    if (gs[i] > 0) std::fprintf(fp, "%d,%d,%llu\n", i, gs[i], (unsigned long long)ps[i]);
            (void)rows;
            std::fclose(fp);
            std::printf("OK gap_sequence.csv\n");
        }
    }
    const unsigned int lgc = (unsigned int)r->large_count;
    std::printf("\n=== Large Gaps (>= %d) ===\nCount: %u\n", 500, lgc);
    {
        const uint64_t tr = std::min<uint64_t>(r->large_count, cap);
        fp = std::fopen("large_gaps.csv", "w");
        if (fp) {
            std::fprintf(fp, "position,gap,merit\n");
            for (uint64_t i = 0; i < tr; i++)
                std::fprintf(fp, "%llu,%d,%.6f\n", (unsigned long long)large[i].position, large[i].gap,
                             large[i].merit);
            std::fclose(fp);
            std::printf("OK large_gaps.csv (%llu entries)\n", (unsigned long long)tr);
            for (uint64_t i = 0; i < tr && i < 10; i++)
                std::printf("  gap=%4d merit=%6.2f at p=%llu\n", large[i].gap, large[i].merit,
                            (unsigned long long)large[i].position);
        }
    }
    std::free(r);
    voss_destroy(ctx);
    std::printf("\nDone.\n");
    return 0;
}
