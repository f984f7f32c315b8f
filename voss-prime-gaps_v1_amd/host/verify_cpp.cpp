// verify_cpp -- large-gap verifier of the drop-in, a host program over
// libvoss_amd.
//
//   verify_cpp [gaps.csv] [details.csv] [cache.csv]
//
// The interface is the one run_voss.py relies on when it runs bin/verify_cpp
// with three paths (run_voss.py:84-87, :218-239; the reference program is
// src/verify_cpp.cpp):
//   * argv: input list, details file, cache file; defaults
//     results/large_gaps.csv, results/verification_details.csv,
//     results/verification_cache.csv;
//   * input: a header line, then rows "p_after,gap[,merit]"; empty rows and
//     rows without a comma are ignored; a missing merit reads as 0;
//   * cache: "p_after,gap" rows after a header; a listed pair counts as
//     verified without being checked again; pairs verified by this run are
//     appended (the header first when the file is new);
//   * details: "position,gap,merit,verified" and one row per input row, in
//     input order, merit printed with %.6f;
//   * exit code 1 when the input cannot be opened or holds no row, or when
//     any row fails; 0 otherwise.
// A row that does not parse as numbers ends the program with exit code 1
// and a message (the reference program aborts there).
//
// The checking itself (verify_gap, src/verify_cpp.cpp:59-93: both ends prime
// and no prime strictly between them) is one voss_verify_gaps call on the
// GPU for every row the cache does not already hold.
#include <algorithm>
#include <cerrno>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "voss/voss.h"

namespace {

struct Row {
    uint64_t p_after;
    int32_t gap;
    double merit;
};

using Key = std::pair<uint64_t, int32_t>;

// Whole file into memory; false if it cannot be opened.
bool slurp(const char* path, std::string& text) {
    FILE* f = std::fopen(path, "rb");
    if (!f) return false;
    char chunk[1 << 16];
    size_t got;
    while ((got = std::fread(chunk, 1, sizeof(chunk), f)) > 0) text.append(chunk, got);
    std::fclose(f);
    return true;
}

// Calls fn(line) for every line after the first (the header).
template <typename F>
void each_data_line(const std::string& text, F&& fn) {
    size_t at = text.find('\n');
    while (at != std::string::npos && at + 1 <= text.size()) {
        const size_t start = at + 1;
        size_t end = text.find('\n', start);
        const size_t stop = end == std::string::npos ? text.size() : end;
        std::string line = text.substr(start, stop - start);
        if (!line.empty() && line.back() == '\r') line.pop_back();
        fn(line);
        at = end;
    }
}

bool to_u64(const char* s, const char* e, uint64_t* out) {
    std::string t(s, e);
    if (t.empty()) return false;
    char* end = nullptr;
    errno = 0;
    const unsigned long long v = std::strtoull(t.c_str(), &end, 10);
    if (errno || end == t.c_str()) return false;
    *out = v;
    return true;
}

bool to_i32(const char* s, const char* e, int32_t* out) {
    std::string t(s, e);
    char* end = nullptr;
    errno = 0;
    const long v = std::strtol(t.c_str(), &end, 10);
    if (errno || end == t.c_str() || v < INT32_MIN || v > INT32_MAX) return false;
    *out = (int32_t)v;
    return true;
}

// "p,g[,merit]" -> row; returns 0 (row), 1 (ignored line), -1 (bad numbers).
int parse_row(const std::string& line, Row* r) {
    if (line.empty()) return 1;
    const size_t c1 = line.find(',');
    if (c1 == std::string::npos) return 1;
    const size_t c2 = line.find(',', c1 + 1);
    const char* s = line.c_str();
    const char* gap_end = s + (c2 == std::string::npos ? line.size() : c2);
    if (!to_u64(s, s + c1, &r->p_after) || !to_i32(s + c1 + 1, gap_end, &r->gap)) return -1;
    r->merit = 0.0;
    if (c2 != std::string::npos) {
        char* end = nullptr;
        r->merit = std::strtod(s + c2 + 1, &end);
        if (end == s + c2 + 1) return -1;
    }
    return 0;
}

// The cache as a sorted list of pairs (binary search per input row).
std::vector<Key> load_cache(const char* path, bool* present) {
    std::vector<Key> keys;
    std::string text;
    *present = slurp(path, text);
    if (!*present) return keys;
    each_data_line(text, [&](const std::string& line) {
        Row r;
        if (parse_row(line, &r) == 0) keys.push_back({r.p_after, r.gap});
    });
    std::sort(keys.begin(), keys.end());
    return keys;
}

}  // namespace

int main(int argc, char** argv) {
    std::setvbuf(stdout, nullptr, _IONBF, 0);
    const char* in_path = argc > 1 ? argv[1] : "results/large_gaps.csv";
    const char* details_path = argc > 2 ? argv[2] : "results/verification_details.csv";
    const char* cache_path = argc > 3 ? argv[3] : "results/verification_cache.csv";
    const auto started = std::chrono::steady_clock::now();

    std::string text;
    if (!slurp(in_path, text)) {
        std::fprintf(stderr, "cannot open %s\n", in_path);
        return 1;
    }
    std::vector<Row> rows;
    bool bad = false;
    each_data_line(text, [&](const std::string& line) {
        Row r;
        const int k = parse_row(line, &r);
        if (k == 0) rows.push_back(r);
        if (k < 0 && !bad) {
            std::fprintf(stderr, "not a gap row in %s: %s\n", in_path, line.c_str());
            bad = true;
        }
    });
    if (bad) return 1;
    std::printf("%zu gap rows in %s\n", rows.size(), in_path);
    if (rows.empty()) return 1;

    bool cache_present = false;
    const std::vector<Key> cache = load_cache(cache_path, &cache_present);
    if (cache_present)
        std::printf("cache %s: %zu verified pairs\n", cache_path, cache.size());
    else
        std::printf("cache %s: none yet, it will be created\n", cache_path);

    // rows the cache already holds are verified; the others go to the GPU
    std::vector<uint8_t> verified(rows.size(), 0);
    std::vector<size_t> pending;
    for (size_t i = 0; i < rows.size(); i++) {
        if (std::binary_search(cache.begin(), cache.end(), Key{rows[i].p_after, rows[i].gap}))
            verified[i] = 1;
        else
            pending.push_back(i);
    }
    std::printf("%zu rows from the cache, %zu to check\n", rows.size() - pending.size(),
                pending.size());

    char device[256] = "";
    if (!pending.empty()) {
        voss_ctx* ctx = nullptr;
        if (voss_create(0, &ctx) != VOSS_OK) {
            std::fprintf(stderr, "VOSS error: %s\n", voss_last_error());
            return EXIT_FAILURE;
        }
        int cus = 0;
        voss_device_name(ctx, device, sizeof(device), &cus, nullptr);
        std::printf("checking on %s (%d CUs)\n", device, cus);
        std::vector<uint64_t> p_after;
        std::vector<int32_t> gap;
        p_after.reserve(pending.size());
        gap.reserve(pending.size());
        for (size_t i : pending) {
            p_after.push_back(rows[i].p_after);
            gap.push_back(rows[i].gap);
        }
        std::vector<uint8_t> ok(pending.size(), 0);
        const int rc = voss_verify_gaps(ctx, p_after.data(), gap.data(), pending.size(), ok.data());
        if (rc != VOSS_OK) {
            std::fprintf(stderr, "VOSS error: %s\n", voss_last_error());
            voss_destroy(ctx);
            return EXIT_FAILURE;
        }
        voss_destroy(ctx);
        for (size_t k = 0; k < pending.size(); k++) verified[pending[k]] = ok[k];
    }

    // newly verified pairs extend the cache
    size_t added = 0;
    if (FILE* cf = std::fopen(cache_path, "a")) {
        if (!cache_present) std::fputs("p_after,gap\n", cf);
        for (size_t i : pending)
            if (verified[i]) {
                std::fprintf(cf, "%llu,%d\n", (unsigned long long)rows[i].p_after, rows[i].gap);
                added++;
            }
        std::fclose(cf);
    }
    size_t failed = 0;
    if (FILE* df = std::fopen(details_path, "w")) {
        std::fputs("position,gap,merit,verified\n", df);
        for (size_t i = 0; i < rows.size(); i++)
            std::fprintf(df, "%llu,%d,%.6f,%d\n", (unsigned long long)rows[i].p_after, rows[i].gap,
                         rows[i].merit, (int)verified[i]);
        std::fclose(df);
    }
    for (uint8_t v : verified) failed += v ? 0 : 1;
    const double secs =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - started).count();
    std::printf("\nverification of %s\n", in_path);
    std::printf("  rows      %zu (%zu from the cache, %zu checked%s%s)\n", rows.size(),
                rows.size() - pending.size(), pending.size(), device[0] ? " on " : "", device);
    std::printf("  verified  %zu\n", rows.size() - failed);
    std::printf("  failed    %zu\n", failed);
    std::printf("  cache     %s (+%zu pairs)\n", cache_path, added);
    std::printf("  details   %s\n", details_path);
    std::printf("  time      %.2f s\n", secs);
    return failed ? 1 : 0;
}
