// call_gate.h -- admission of library calls: one thread inside the library
// at a time per process.
//
// Calls on two contexts from two host threads at once are not supported:
// while one thread's stream is between hipStreamBeginCapture and
// hipStreamEndCapture, or inside hipGraphInstantiate, the other thread's
// calls may synchronise the device (hipFree, hipHostFree, hipMalloc,
// null-stream hipMemcpy, hipMemcpyToSymbol), and one such run stopped making
// progress.  The gate turns that use into an error: an entry point enters
// with try_enter() and returns VOSS_EBUSY when another thread is inside;
// setup and teardown calls and the serving scheduler's steps wait with
// enter().  A thread that holds the gate may enter again (the library's
// entry points call each other), so the count is per thread.
#pragma once
#include <mutex>

namespace voss {

class CallGate {
public:
    // true: the calling thread is now inside (or already was)
    bool try_enter() {
        if (depth_ > 0) {
            depth_++;
            return true;
        }
        if (!mu_.try_lock()) return false;
        depth_ = 1;
        return true;
    }
    void enter() {
        if (depth_ > 0) {
            depth_++;
            return;
        }
        mu_.lock();
        depth_ = 1;
    }
    // once per successful try_enter() / enter()
    void leave() {
        if (--depth_ == 0) mu_.unlock();
    }
    bool held_by_caller() const { return depth_ > 0; }

private:
    std::mutex mu_;
    // entries of the calling thread (a thread holds at most one gate: the
    // library has a single process-wide instance)
    static inline thread_local int depth_ = 0;
};

}  // namespace voss
