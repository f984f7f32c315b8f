// serve_launch.h -- the device-resident serving loop (serve_kernel.hip):
// records shared by the kernel and the host scheduler (serve.hip).
//
// The host publishes units of work into a ring in pinned host memory and
// advances `tail`; the persistent kernel's workgroups claim units in ring
// order (a compare-and-swap on a device counter, never past `tail`), so a
// unit is claimed only after every unit before it.  Each workgroup reports
// the last unit it claimed, so the host keeps only a few units per
// workgroup published ahead of the device and chooses the next ones from
// all active requests in turn (a request that arrives while a long one runs
// gets its units in within a few windows' time).  A request's bucket
// tasks precede its windows in the ring, so a window that waits for its
// tile's bucket tasks only waits for tasks that running workgroups hold.
// Per-request state that the device reads (descriptors, bucket tasks and
// offsets, the large-gap list) is in pinned host memory; state that it
// updates is in HBM.  The workgroup that completes a request's last unit
// joins the pairs across its windows, copies its sums into the request's
// mailbox in pinned host memory, clears the slot's device state for the
// next request and publishes the request's generation.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sieve_launch.h"
#include "wheel.h"

namespace voss {

enum ServeUnitKind : uint32_t {
    kUnitWindow = 1,  // index = window of an interval request
    kUnitBucket = 2,  // index = bucket task of an interval request
    kUnitMr = 3,      // index = chunk of kServeMrChunk candidates of an MR request
};

constexpr uint32_t kServeMrChunk = 8192;

struct ServeUnit {
    uint32_t kind;
    uint32_t slot;  // interval slot, or MR slot
    uint32_t index;
    uint32_t gen;   // generation of the request in that slot (diagnostics)
    unsigned long long seq;  // the unit's position in the ring sequence: the
                             // device re-reads an entry that does not carry it yet
};

// An interval request in its slot (host memory, written before its units
// are published).  Bucket indices are the request's own (0 .. nbkt-1);
// tile t holds buckets [256 t, 256 t + 256).
struct ServeReqInfo {
    const SegDesc* desc;               // [nwin]
    const BucketTask* tasks;           // [ntask]
    const unsigned long long* boff;    // [nbkt + 1] first entry of each bucket in the arena
    const uint32_t* tile_ntask;        // [ntile] bucket tasks of each tile
    uint32_t* tile_ready;              // host [ntile]: set when the tile's tasks are done
    uint32_t* cursor;                  // device [nbkt], zero while free
    uint32_t* tile_done;               // device [ntile], zero while free
    SegRec* rec;                       // device [nwin]
    LargeRaw* large;                   // host [lcap]
    unsigned long long lcap;
    uint32_t nwin, ntask, nbkt, ntile;
    uint32_t gen, pad;
};

struct ServeMrInfo {
    const uint64_t* cand;   // host [count]
    uint8_t* verdict;       // host [count], written in 4-byte words: rounded up to 4
    unsigned long long count;
    uint32_t nchunk;
    uint32_t gen;
};

// Device state of an interval slot, zero while the slot is free.
struct ServeSlot {
    unsigned long long large_n;  // large-gap candidates so far
    uint32_t windows_done;
    uint32_t error;              // bit 0: a bucket dependency wait gave up
    uint32_t overflow;           // a bucket outgrew its room
    uint32_t pad;
};

// Host-memory results of an interval request, valid once done_gen equals
// the request's generation.
struct ServeMailbox {
    unsigned long long done_gen;
    unsigned long long large_n;
    uint32_t error;
    uint32_t overflow;
    JobAcc acc;
};

struct ServeCtl {
    unsigned long long tail;  // units published (host writes)
    uint32_t stop;            // 1: exit when idle, 2: exit at the next claim
    uint32_t fault;           // device: a claimed ring entry never showed its stamp
};

struct ServeArgs {
    const ServeUnit* ring;         // host [ring_mask + 1]
    unsigned long long ring_mask;
    ServeCtl* ctl;                 // host
    unsigned long long* claim;     // device: [0] units claimed so far, [16] the
                                   // largest tail seen, [17] stop seen, [32] the
                                   // poll ticket (one workgroup reads the host's
                                   // tail and stop at a time); on separate lines
    const ServeReqInfo* req;       // host [max_active]
    const ServeMrInfo* mrq;        // host [mr_slots]
    ServeMailbox* mail;            // host [max_active]
    unsigned long long* mr_mail;   // host [mr_slots]: done generation
    ServeSlot* slot;               // device [max_active]
    uint32_t* mr_done;             // device [mr_slots]: chunks completed, zero while free
    JobAcc* acc;                   // device [max_active], initial state while free
    uint32_t* entries;             // device bucket-entry arena
    unsigned long long* units_done;  // device: units completed (statistics)
    unsigned long long* claimed_by;  // host [grid]: per workgroup, 1 + the last unit it
                                     // claimed (units are claimed in order, so the
                                     // largest is a lower bound of the claim count)
    unsigned long long* wg_stats;    // host [grid][8]: per workgroup and unit kind k,
                                     // [2k] units run, [2k+1] wall-clock ticks spent
    uint32_t idle_limit;           // empty polls before a workgroup exits
    uint32_t dep_limit;            // polls of a bucket dependency before giving up
    int log2w;                     // log2 of the window bytes
};

// Launches the persistent loop (grid = one workgroup per CU).  The kernel
// returns when every workgroup has found the ring empty idle_limit times in
// a row, or has seen ctl->stop.
hipError_t launch_serve(const SieveArgs& a, const ServeArgs& sv, uint32_t grid, hipStream_t s);

}  // namespace voss
