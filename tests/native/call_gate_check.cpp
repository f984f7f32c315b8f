// Host-only checks of voss::CallGate (csrc/call_gate.h): exclusion between
// threads, re-entry by the holding thread, and that enter() waits for the
// holder instead of failing.  Prints "ok" and exits 0 when every check holds.
#include <atomic>
#include <chrono>
#include <cstdio>
#include <thread>

#include "call_gate.h"

#define CHECK(c)                                                        \
    do {                                                                \
        if (!(c)) {                                                     \
            std::fprintf(stderr, "check failed line %d: %s\n", __LINE__, #c); \
            return 1;                                                   \
        }                                                               \
    } while (0)

int main() {
    voss::CallGate gate;
    // re-entry by the holder; released only after the last leave
    CHECK(gate.try_enter());
    CHECK(gate.try_enter());
    gate.enter();
    CHECK(gate.held_by_caller());
    std::atomic<int> other{-1};
    std::thread([&] { other = gate.try_enter() ? 1 : 0; if (other == 1) gate.leave(); }).join();
    CHECK(other == 0);  // held: another thread is refused
    gate.leave();
    gate.leave();
    std::thread([&] { other = gate.try_enter() ? 1 : 0; if (other == 1) gate.leave(); }).join();
    CHECK(other == 0);  // one entry left
    gate.leave();
    CHECK(!gate.held_by_caller());
    std::thread([&] { other = gate.try_enter() ? 1 : 0; if (other == 1) gate.leave(); }).join();
    CHECK(other == 1);  // free again

    // enter() waits for the holder
    std::atomic<bool> holding{false}, released{false}, got{false};
    std::thread holder([&] {
        gate.enter();
        holding = true;
        std::this_thread::sleep_for(std::chrono::milliseconds(50));
        released = true;
        gate.leave();
    });
    while (!holding) std::this_thread::yield();
    std::thread waiter([&] {
        gate.enter();
        got = released.load();  // must only get in after the release
        gate.leave();
    });
    holder.join();
    waiter.join();
    CHECK(got);

    // many threads: the section never holds two threads at once
    std::atomic<int> inside{0}, peak{0}, refused{0}, entered{0};
    std::thread ts[8];
    for (int t = 0; t < 8; t++)
        ts[t] = std::thread([&, t] {
            for (int k = 0; k < 2000; k++) {
                const bool ok = (t & 1) ? gate.try_enter() : (gate.enter(), true);
                if (!ok) { refused++; continue; }
                const int now = ++inside;
                int p = peak.load();
                while (now > p && !peak.compare_exchange_weak(p, now)) {}
                if (gate.try_enter()) gate.leave();  // nested entry always passes
                --inside;
                entered++;
                gate.leave();
            }
        });
    for (auto& t : ts) t.join();
    CHECK(peak == 1);
    CHECK(entered + refused == 16000);
    std::printf("ok\n");
    return 0;
}
