// verify_cpp -- the large-gap verifier of the drop-in, over libvoss_amd.
//
//   verify_cpp [large_gaps.csv] [verification_details.csv] [verification_cache.csv]
//
// Same command line, files, cache semantics and exit code as the
// reference's src/verify_cpp.cpp main() (:112-276), which run_voss.py runs
// as bin/verify_cpp with those three paths (run_voss.py:84-87, :218-239).
// The per-gap work -- Miller-Rabin on both ends and the proof that every
// number in between is composite (verify_gap, :59-93) -- runs on the GPU:
// 12-base Montgomery Miller-Rabin and the LDS-window interval sieve.
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <unordered_set>
#include <utility>
#include <vector>

#include "voss/voss.h"

namespace {

struct PairHash {
    size_t operator()(const std::pair<uint64_t, int>& p) const {
        return std::hash<uint64_t>()(p.first) * 0x9e3779b97f4a7c15ULL ^ std::hash<int>()(p.second);
    }
};

}  // namespace

int main(int argc, char** argv) {
    setvbuf(stdout, nullptr, _IONBF, 0);
    const char* input = (argc > 1) ? argv[1] : "results/large_gaps.csv";
    const char* output = (argc > 2) ? argv[2] : "results/verification_details.csv";
    const char* cache_path = (argc > 3) ? argv[3] : "results/verification_cache.csv";
    auto t0 = std::chrono::steady_clock::now();

    std::ifstream fin(input);
    if (!fin) {
        std::fprintf(stderr, "cannot open %s\n", input);
        return 1;
    }
    std::vector<std::pair<uint64_t, int>> gaps;
    std::vector<double> merits;
    std::string line;
    std::getline(fin, line);  // header
    while (std::getline(fin, line)) {
        if (line.empty()) continue;
        const size_t c1 = line.find(',');
        if (c1 == std::string::npos) continue;
        const size_t c2 = line.find(',', c1 + 1);
        try {
            const uint64_t p = std::stoull(line.substr(0, c1));
            const int g = std::stoi(line.substr(c1 + 1, c2 - c1 - 1));
            const double m = (c2 != std::string::npos) ? std::stod(line.substr(c2 + 1)) : 0.0;
            gaps.push_back({p, g});
            merits.push_back(m);
        } catch (const std::exception&) {
            std::fprintf(stderr, "bad row in %s: %s\n", input, line.c_str());
            return 1;
        }
    }
    fin.close();
    const int n = (int)gaps.size();
    std::printf("Read %d gaps from %s\n", n, input);
    if (n == 0) return 1;

    std::unordered_set<std::pair<uint64_t, int>, PairHash> cache;
    {
        std::ifstream cf(cache_path);
        if (cf) {
            std::string cl;
            std::getline(cf, cl);
            while (std::getline(cf, cl)) {
                if (cl.empty()) continue;
                const size_t c = cl.find(',');
                if (c == std::string::npos) continue;
                cache.insert({std::stoull(cl.substr(0, c)), std::stoi(cl.substr(c + 1))});
            }
            std::printf("Loaded %zu cached verifications from %s\n", cache.size(), cache_path);
        } else {
            std::printf("No cache yet - will create: %s\n", cache_path);
        }
    }
    std::vector<int> to_verify;
    std::vector<uint8_t> ok(n, 0);
    int cached_count = 0;
    for (int i = 0; i < n; i++) {
        if (cache.count(gaps[i])) {
            ok[i] = 1;
            cached_count++;
        } else {
            to_verify.push_back(i);
        }
    }
    std::printf("Cached: %d, To verify: %zu\n", cached_count, to_verify.size());

    voss_ctx* ctx = nullptr;
    if (voss_create(0, &ctx) != VOSS_OK) {
        std::fprintf(stderr, "VOSS error: %s\n", voss_last_error());
        return EXIT_FAILURE;
    }
    char name[256];
    int cus = 0;
    voss_device_name(ctx, name, sizeof(name), &cus, nullptr);
    std::printf("Verifying on %s (%d CUs)...\n", name, cus);
    if (!to_verify.empty()) {
        std::vector<uint64_t> pa(to_verify.size());
        std::vector<int32_t> g(to_verify.size());
        std::vector<uint8_t> res(to_verify.size());
        for (size_t k = 0; k < to_verify.size(); k++) {
            pa[k] = gaps[to_verify[k]].first;
            g[k] = gaps[to_verify[k]].second;
        }
        if (voss_verify_gaps(ctx, pa.data(), g.data(), pa.size(), res.data()) != VOSS_OK) {
            std::fprintf(stderr, "VOSS error: %s\n", voss_last_error());
            voss_destroy(ctx);
            return EXIT_FAILURE;
        }
        for (size_t k = 0; k < to_verify.size(); k++) ok[to_verify[k]] = res[k];
    }
    voss_destroy(ctx);

    long long verified = 0, failed = 0;
    for (int i = 0; i < n; i++) (ok[i] ? verified : failed)++;
    {
        std::ifstream check(cache_path);
        const bool exists = check.good();
        check.close();
        std::ofstream cf(cache_path, std::ios::app);
        if (!exists) cf << "p_after,gap\n";
        size_t appended = 0;
        for (int i : to_verify)
            if (ok[i]) {
                cf << gaps[i].first << "," << gaps[i].second << "\n";
                appended++;
            }
        std::printf("Appended %zu new verifications to cache\n", appended);
    }
    std::ofstream fout(output);
    fout << "position,gap,merit,verified\n";
    char buf[160];
    for (int i = 0; i < n; i++) {
        std::snprintf(buf, sizeof(buf), "%llu,%d,%.6f,%d", (unsigned long long)gaps[i].first,
                      gaps[i].second, merits[i], (int)ok[i]);
        fout << buf << "\n";
    }
    fout.close();
    const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::printf("\n============================================================\n");
    std::printf("GPU VERIFIER RESULT (cached)\n");
    std::printf("============================================================\n");
    std::printf("Total gaps:    %d\n", n);
    std::printf("Cached:        %d  (%.1f%%)\n", cached_count, 100.0 * cached_count / n);
    std::printf("Newly verify:  %zu\n", to_verify.size());
    std::printf("Verified:      %lld\n", verified);
    std::printf("Failed:        %lld\n", failed);
    std::printf("Device:        %s\n", name);
    std::printf("Duration:      %.2f s\n", dt);
    std::printf("Details:       %s\n", output);
    std::printf("Cache:         %s\n", cache_path);
    std::printf("============================================================\n");
    return (failed > 0) ? 1 : 0;
}
