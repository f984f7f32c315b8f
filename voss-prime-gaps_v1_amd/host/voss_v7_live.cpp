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
#include <cstring>
#include <string>
#include <vector>

#include "voss/voss.h"

namespace {

constexpr int64_t kMaxGap = VOSS_MAX_GAP;
constexpr int64_t kMaxLarge = 50000000;  // MAX_LARGE_GAPS (:34)

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

void fmt_time(double s, char* buf) {
    if (s < 60.0) std::snprintf(buf, 32, "%5.1f s", s);
    else if (s < 3600.0) std::snprintf(buf, 32, "%5.1f m", s / 60.0);
    else std::snprintf(buf, 32, "%5.1f h", s / 3600.0);
}

// print_progress (:317-335)
void print_progress(int64_t seg, int64_t num_seg, uint64_t total_primes, double elapsed_s,
                    uint64_t last_prime) {
    int percent = (int)((seg * 100) / num_seg);
    double rate = (elapsed_s > 0.1) ? ((double)seg / elapsed_s) : 0.0;
    double eta_s = (rate > 0.0) ? ((double)(num_seg - seg) / rate) : 0.0;
    char eta_str[32], el_str[32];
    fmt_time(eta_s, eta_str);
    fmt_time(elapsed_s, el_str);
    std::fprintf(stdout,
                 "\r[%3d%%] seg %lld/%lld | primes: %-13llu | last p: %-16llu | elapsed: %s | ETA: %s   ",
                 percent, (long long)seg, (long long)num_seg, (unsigned long long)total_primes,
                 (unsigned long long)last_prime, el_str, eta_str);
    std::fflush(stdout);
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
    std::printf(">>> STARTING FRESH <<<\n");

    voss_result* r = (voss_result*)std::calloc(1, sizeof(voss_result));
    const uint64_t cap = std::min<uint64_t>(max_large, (uint64_t)kMaxLarge);
    std::vector<voss_large_gap> large(cap ? cap : 1);
    r->large = large.data();
    r->large_capacity = cap;

    auto t0 = std::chrono::steady_clock::now();
    if (voss_run_plain(ctx, N, SEG_NUM, r) != VOSS_OK) die("voss_run_plain");
    const double total_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    print_progress(NUM_SEG, NUM_SEG, r->total_primes, total_s, r->last_prime);
    std::printf("\n\n");
    std::remove("checkpoint.bin");

    int64_t* gaps = r->gaps;
    std::printf("=============================================================\n");
    std::printf("TOTAL TIME: %.2f s (%.2f min)\n", total_s, total_s / 60.0);
    std::printf("Total primes: %llu\n", (unsigned long long)r->total_primes);
    std::printf("=============================================================\n\n");

    std::printf("=== Timing Breakdown ===\n");
    std::printf("Sieve+extract+gaps (fused kernel): %10.3f s\n", r->sieve_ms / 1000.0);
    std::printf("Stitch (window edges):             %10.3f s\n", (r->device_ms - r->sieve_ms) / 1000.0);
    std::printf("Host (plan, copies, finalise):     %10.3f s\n\n", r->seconds - r->device_ms / 1000.0);

    // statistics (:652-672)
    int64_t tgc = 0, wsum = 0;
    for (int g = 1; g < kMaxGap; g++) { tgc += gaps[g]; wsum += (int64_t)g * gaps[g]; }
    double mg = (double)wsum / (double)tgc;
    double vs = 0, ss = 0, ks = 0;
    for (int g = 1; g < kMaxGap; g++) {
        if (gaps[g] == 0) continue;
        double d = (double)g - mg;
        vs += (double)gaps[g] * d * d;
        ss += (double)gaps[g] * d * d * d;
        ks += (double)gaps[g] * d * d * d * d;
    }
    double var = vs / (double)tgc, sd = std::sqrt(var);
    double sk = ss / (double)tgc / (var * sd);
    double ku = ks / (double)tgc / (var * var) - 3.0;
    std::printf("=== Gap Statistics ===\n");
    std::printf("Total gaps:  %lld\n", (long long)tgc);
    std::printf("Mean gap:    %.4f\n", mg);
    std::printf("Std dev:     %.4f\n", sd);
    std::printf("Skewness:    %.4f\n", sk);
    std::printf("Kurtosis:    %.4f\n\n", ku);
    std::printf("=== Gap Counts ===\n");
    for (int g : {1, 2, 4, 6, 12, 30})
        std::printf("Gap %d:%s%lld\n", g, g < 10 ? "  " : " ", (long long)gaps[g]);
    std::printf("\n");

    const unsigned long long p41 = r->c1 + 1, p43 = r->c3 + 1;  // :682
    const long long diff = (long long)p43 - (long long)p41;
    std::printf("=== Chebyshev Bias ===\n");
    std::printf("pi(x;4,1) = %llu\n", p41);
    std::printf("pi(x;4,3) = %llu\n", p43);
    std::printf("Diff      = %lld\n\n", diff);
    const double p6p2 = (double)gaps[6] / (double)gaps[2];
    std::printf("=== Gap Repulsion ===\n");
    std::printf("P(6)/P(2) = %.6f\n\n", p6p2);

    const uint64_t total = r->total_primes;
    std::printf("=== Self-Consistency Checks ===\n");
    std::printf("[Check 1] sum(N(g)) = pi(N) - 1\n");
    std::printf("  sum(N(g))  = %lld\n", (long long)tgc);
    std::printf("  pi(N) - 1  = %lld\n", (long long)(total - 1));
    std::printf("  -> %s\n", (tgc == (int64_t)(total - 1)) ? "PASS" : "FAIL");
    std::printf("[Check 2] sum(g*N(g)) = p_last - 2\n");
    std::printf("  sum(g*N(g)) = %lld\n", (long long)wsum);
    std::printf("  p_last - 2  = %lld\n", (long long)(r->last_prime - 2));
    std::printf("  -> %s\n", (wsum == (int64_t)(r->last_prime - 2)) ? "PASS" : "FAIL");
    std::printf("[Check 3] pi(4,1) + pi(4,3) + 1 = pi(N)\n");
    std::printf("  pi41+pi43+1 = %llu\n", p41 + p43 + 1);
    std::printf("  pi(N)       = %llu\n", (unsigned long long)total);
    std::printf("  -> %s\n\n", ((p41 + p43 + 1) == total) ? "PASS" : "FAIL");

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
        std::fprintf(fp, "%lld,%llu,%llu,%lld\n", (long long)N, p41, p43, diff);
        std::fclose(fp);
        std::printf("OK chebyshev.csv\n");
    }
    {
        std::vector<int32_t> gs(VOSS_GAP_SEQ_SIZE, 0);
        std::vector<uint64_t> ps(VOSS_GAP_SEQ_SIZE, 0);
        uint64_t rows = 0;
        const int rc = voss_sample_plain(ctx, N, SEG_NUM, gs.data(), ps.data(), &rows);
        if (rc != VOSS_OK && rc != VOSS_EINVAL) die("voss_sample_plain");
        fp = std::fopen("gap_sequence.csv", "w");
        if (fp) {
            std::fprintf(fp, "index,gap,p\n");
            if (rc == VOSS_OK)
                for (int i = 0; i < VOSS_GAP_SEQ_SIZE; i++)
                    if (gs[i] > 0) std::fprintf(fp, "%d,%d,%llu\n", i, gs[i], (unsigned long long)ps[i]);
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
