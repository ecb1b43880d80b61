// rt_serve_core.hpp — the serving loop's request states, admission policy and
// memory accounting, with no device calls.
//
// rt_server.cpp drives one Core from its worker thread (expire, admit,
// launch, finished) and from the client entry points (submit, cancel, poll,
// wait, take). The CPU tests drive the same Core with a fake clock and a fake
// device, under ThreadSanitizer and AddressSanitizer.
//
// Three resources are admitted against: device memory, pinned staging and the
// device's wave slots (workgroups that can be resident at once). Every kernel
// of the serving loop is sized to the slots it was granted, and grants never
// exceed the free slots, so all of a kernel's workgroups are resident from the
// start and the hardware never holds one request's kernel back behind
// another's undispatched workgroups (measured: a 3 ms request launched next to
// a larger-than-free superposition waited out its whole 136 ms). A
// single-Kerr request gets the free slots up to its need. A superposition gets
// at most three quarters of all slots and never the last quarter
// (multi_reserve): its rays take milliseconds to hundreds of them, and
// single-Kerr requests arriving meanwhile must still find room. (Measured on
// the serving mix: three quarters against half cut the p99 of all requests
// from 137 to 121 ms at 50 QPS and from 116 to 98 ms at 25 QPS, unchanged at
// 100 QPS; seven eighths with an eighth kept back doubled it at 100 QPS.) A request whose grant would be
// below min(its need, min_grant) waits for more room instead (a request on a
// sliver of the device would hold its lane and memory for a long time).
// Slots come back when the request's kernel ends. Slots are never taken back from a running kernel: its waves
// could only leave between rays, which for a superposition can take a long
// time, and until then the new kernel would not be resident. A cancelled or
// expired running request is asked to stop early (Resize to 0: its waves
// leave between tiles and its results are dropped); its slots come back when
// its kernel ends. A request also needs a lane of its own (a stream, one per
// hardware queue): kernels on one stream run one after another, so a granted
// request must not queue behind another. The last kerr_only_lanes lanes take
// single-Kerr requests only.
//
// Life of a request:
//   submit -> Pending -> (admit: device and staging blocks, slots, a lane) Running
//          -> (finished) Done -> take_begin / take_end (staging block freed)
//   Pending  + cancel               -> Cancelled
//   Pending  + deadline passed      -> Expired (RT_ERR_EXPIRED)
//   Running  + cancel or deadline   -> Cancelled / Expired at once; its blocks
//                                      stay held until the device finishes
//   Done     + cancel               -> Cancelled (staging block freed)
//   Running  + launch or device err -> Failed
// A finished state is reported by poll/wait; take hands back Done results or
// the final status and forgets the request once the device no longer uses it.
//
// The windowed worker it replaces holds one pending frame and drops older ones
// (render/geodesic_compute_pipeline.hpp:529-728); deadlines follow the
// scheduler's in-flight expiry (src/distributed/distribution_scheduler.cpp:389-405).
#pragma once
#include <condition_variable>
#include <cstddef>
#include <cstdint>
#include <deque>
#include <functional>
#include <map>
#include <mutex>
#include <vector>

#include "../../include/rt_tracer.h"

namespace rt {
namespace serve {

// Monotonic time in microseconds. The server uses steady_now_us; tests
// inject their own.
using Clock = std::function<int64_t()>;
int64_t steady_now_us();

// First-fit allocator over [0, size) in 256-byte units.
class PoolAlloc {
public:
    void init(size_t size);
    bool alloc(size_t bytes, size_t* off);
    void release(size_t off, size_t bytes);
    size_t free_bytes() const;
    size_t size() const { return size_; }

private:
    std::map<size_t, size_t> free_;  // offset -> bytes
    size_t size_ = 0;
};

enum class State : uint8_t { Pending, Running, Done, Cancelled, Expired, Failed };

// What the core needs to know of a request.
struct Job {
    size_t dev_bytes = 0;     // device block: results, tables, counters
    size_t host_bytes = 0;    // pinned staging block: the outputs asked for (+ upload tables)
    bool multi = false;       // superposition: not on the single-Kerr-only lanes
    int64_t deadline_us = 0;  // absolute, on the core's clock; 0: none
    uint32_t slots_need = 1;  // wave slots it can use at most (>= 1)
};

struct Admission {
    uint64_t id = 0;
    size_t dev_off = 0, host_off = 0;
    uint32_t lane = 0;
    uint32_t slots = 0;  // wave slots granted: the kernel's workgroup count
};

// A running request's new wave limit (0: stop early, results dropped).
struct Resize {
    uint64_t id = 0;
    uint32_t slots = 0;
};

struct Config {
    size_t dev_bytes = 0;        // device pool
    size_t host_bytes = 0;       // pinned staging pool
    uint32_t max_inflight = 64;  // requests on the device at once (admitted, not finished)
    uint32_t lanes = 4;          // streams; each runs one request at a time
    uint32_t kerr_only_lanes = 1;  // the last ones take single-Kerr requests only
    int64_t starve_us = 20000;   // a pending request older than this is not overtaken
    uint32_t slots = 0;          // the device's wave slots (0: not limited)
    uint32_t multi_share = 0;    // most slots one superposition gets (0: slots * 3 / 4)
    uint32_t multi_reserve = 0;  // slots superpositions leave free (0: slots / 4)
    uint32_t min_grant = 0;      // smallest grant worth starting (0: slots / 8)
};

struct Counts {
    size_t pending = 0, running = 0, finished = 0;  // finished: results or status not yet taken
    size_t dev_free = 0, host_free = 0;
    uint32_t slots_free = 0;
};

class Core {
public:
    Core(const Config& cfg, Clock clock);

    // ---- client side (any thread)
    // RT_ERR_BUSY when the request can never fit the pools.
    rt_status submit(const Job& job, uint64_t* id);
    rt_status cancel(uint64_t id);
    rt_status poll(uint64_t id, int* done);
    // Wait (real time) until the request is finished; RT_ERR_TIMEOUT when
    // timeout_ms (>= 0) passes first. Returns the request's final status.
    rt_status wait(uint64_t id, int64_t timeout_ms);
    // Done: RT_OK with the staging offset; the request is then Taking until
    // take_end. Otherwise its final status (RT_ERR_BUSY while unfinished); a
    // request that is not Done is forgotten here (or when the device lets go).
    rt_status take_begin(uint64_t id, size_t* host_off);
    // give_back: the caller could not take the results; the request stays Done.
    void take_end(uint64_t id, bool give_back = false);

    // ---- worker side
    // Requests whose deadline passed: pending ones end Expired; running ones
    // are reported Expired and their results dropped when they finish.
    void expire(std::vector<uint64_t>* dropped = nullptr);
    // Admit pending requests in arrival order while blocks and slots allow.
    // resized receives the wave limits of running requests to lower (0 after
    // a cancel or expiry), in order.
    // A request that does not fit is passed by later ones until it has waited
    // starve_us, then nothing passes it.
    void admit(std::vector<Admission>* out, std::vector<Resize>* resized = nullptr);
    // The launch of an admitted request failed: Failed, blocks freed.
    void launch_failed(uint64_t id, rt_status st);
    // Kernel and copies of a running request are complete (st != RT_OK: the
    // device reported an error). Frees the device block.
    void finished(uint64_t id, rt_status st);
    // Block the worker until new work may be admitted or timeout_us passes.
    void wait_work(int64_t timeout_us);
    // Wake the worker (shutdown).
    void notify_worker();
    // Shut down: pending requests end Cancelled; waiters wake.
    void shutdown();

    Counts counts() const;
    bool find_state(uint64_t id, State* st) const;
    // Lanes running a request.
    uint32_t lanes_busy() const;

private:
    struct Rec {
        Job job;
        State state = State::Pending;
        rt_status err = RT_OK;
        bool on_device = false;  // device block held (admitted, not finished)
        bool has_host = false;   // staging block held
        bool taking = false;
        bool forget = false;     // drop the record once the device lets go
        size_t dev_off = 0, host_off = 0;
        uint32_t lane = 0;
        uint32_t slots = 0;
        int64_t submit_us = 0;
    };
    static bool finished_state(State s) { return s != State::Pending && s != State::Running; }
    void release_host(Rec& r);
    void end_locked(uint64_t id, Rec& r, State s, rt_status err);
    void maybe_erase(std::map<uint64_t, Rec>::iterator it);
    // Grant wave slots to a new request (see the policy above); 0: none now.
    uint32_t take_slots(uint32_t need, bool multi);

    Config cfg_;
    Clock clock_;
    mutable std::mutex mu_;
    std::condition_variable cv_done_, cv_work_;
    bool work_flag_ = false;
    bool stopped_ = false;
    uint64_t next_id_ = 1;
    std::map<uint64_t, Rec> recs_;
    std::deque<uint64_t> pending_;
    PoolAlloc dev_, host_;
    std::vector<bool> lane_busy_;
    std::vector<Resize> resizes_;  // limits to hand to the worker at the next admit
    size_t running_ = 0;       // requests holding device blocks
    uint32_t slots_used_ = 0;  // wave slots granted to running requests
};

}  // namespace serve
}  // namespace rt
