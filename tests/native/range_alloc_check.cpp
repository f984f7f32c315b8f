// Host-only check of the device loop's arena allocator (range_alloc.h)
// against a brute-force occupancy map: random allocations and releases,
// every allocation inside [0, cap), aligned, disjoint from every live one;
// an allocation fails only when no aligned free run is long enough; the
// free list stays merged; after releasing everything one range is left.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "range_alloc.h"

using voss::serving::RangeAlloc;

static int fails = 0;
#define CHECK(c)                                                      \
    do {                                                              \
        if (!(c)) {                                                   \
            std::fprintf(stderr, "line %d: %s\n", __LINE__, #c);      \
            if (++fails > 20) std::exit(1);                           \
        }                                                             \
    } while (0)

// smallest aligned offset with n free units in the map, or -1
static long long brute_first_fit(const std::vector<char>& used, uint64_t n, uint64_t align) {
    const uint64_t cap = used.size();
    for (uint64_t o = 0; o + n <= cap; o += align) {
        bool ok = true;
        for (uint64_t k = 0; k < n && ok; k++) ok = !used[o + k];
        if (ok) return (long long)o;
    }
    return -1;
}

int main() {
    std::mt19937_64 rng(0x564F5353);
    for (int round = 0; round < 40; round++) {
        const uint64_t cap = 1 + rng() % 3000;
        RangeAlloc A;
        A.reset(cap);
        std::vector<char> used(cap, 0);
        struct Live { uint64_t off, n; };
        std::vector<Live> live;
        for (int step = 0; step < 400; step++) {
            if (live.empty() || rng() % 3 != 0) {
                const uint64_t n = rng() % 5 == 0 ? 0 : 1 + rng() % (cap / 4 + 1);
                const uint64_t align = uint64_t(1) << (rng() % 5);
                uint64_t off = ~0ull;
                const bool ok = A.alloc(n, align, &off);
                if (n == 0) {
                    CHECK(ok);
                    continue;
                }
                // first fit over free runs: it may pick a later run than the
                // brute force's smallest offset only if that one does not start
                // a free run of the map, so check feasibility and validity
                const long long bf = brute_first_fit(used, n, align);
                CHECK(ok == (bf >= 0));
                if (!ok) continue;
                CHECK(off % align == 0);
                CHECK(off + n <= cap);
                for (uint64_t k = 0; k < n; k++) {
                    CHECK(!used[off + k]);
                    used[off + k] = 1;
                }
                live.push_back({off, n});
            } else {
                const size_t i = rng() % live.size();
                A.release(live[i].off, live[i].n);
                for (uint64_t k = 0; k < live[i].n; k++) used[live[i].off + k] = 0;
                live[i] = live.back();
                live.pop_back();
            }
            // the free list is the map's free runs, merged
            uint64_t total = 0, prev_end = ~0ull;
            for (const auto& f : A.free_) {
                CHECK(f.second > 0);
                CHECK(f.first != prev_end);  // merged with its left neighbour
                for (uint64_t k = 0; k < f.second; k++) CHECK(!used[f.first + k]);
                total += f.second;
                prev_end = f.first + f.second;
            }
            uint64_t free_units = 0;
            for (char u : used) free_units += !u;
            CHECK(total == free_units);
            CHECK(A.free_total() == free_units);
        }
        for (const Live& l : live) A.release(l.off, l.n);
        CHECK(A.free_.size() == 1 && A.free_.begin()->first == 0 &&
              A.free_.begin()->second == cap);
        CHECK(A.largest() == cap);
    }
    if (fails) return 1;
    std::printf("ok\n");
    return 0;
}
