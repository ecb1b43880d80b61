// serve_core_check.cpp — CPU checks of the serving core (rt_serve_core.cpp):
// allocator, admission by footprint, refusal, overtaking and its age limit,
// cancel and deadline in every state with an injected clock, lanes, the take
// protocol, shutdown, and a multi-threaded run (clients, a fake device, a
// clock thread) whose end state must be empty. Built and run by
// tests/test_serve_core.py under ThreadSanitizer and under AddressSanitizer +
// UndefinedBehaviorSanitizer. Prints one line per check; exits 1 on failure.
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <random>
#include <set>
#include <string>
#include <thread>
#include <vector>

#include "rt_serve_core.hpp"

using namespace rt::serve;

static std::atomic<int> g_fail{0};
#define CHECK(c)                                                             \
    do {                                                                     \
        if (!(c)) {                                                          \
            std::fprintf(stderr, "FAIL %s:%d: %s\n", __FILE__, __LINE__, #c); \
            ++g_fail;                                                        \
        }                                                                    \
    } while (0)

struct FakeClock {
    std::atomic<int64_t> t{1000000};
    Clock fn() {
        return [this] { return t.load(); };
    }
};

constexpr size_t U = 256;  // allocation unit

static Job job(size_t dev_units, size_t host_units = 1, bool multi = false, int64_t deadline = 0) {
    Job j;
    j.dev_bytes = dev_units * U;
    j.host_bytes = host_units * U;
    j.multi = multi;
    j.deadline_us = deadline;
    return j;
}

static Config cfg(size_t dev_units, size_t host_units, uint32_t inflight = 64) {
    Config c;
    c.dev_bytes = dev_units * U;
    c.host_bytes = host_units * U;
    c.max_inflight = inflight;
    c.lanes = 64;  // lanes are checked on their own below
    c.kerr_only_lanes = 0;
    c.starve_us = 20000;
    return c;
}

static Job sjob(uint32_t need, size_t dev_units = 1) {
    Job j = job(dev_units);
    j.slots_need = need;
    return j;
}

static std::vector<Admission> admit(Core& c) {
    std::vector<Admission> a;
    c.admit(&a);
    return a;
}

static State state(Core& c, uint64_t id) {
    State s = State::Failed;
    CHECK(c.find_state(id, &s));
    return s;
}

// ---------------------------------------------------------------- checks
static void pool_alloc() {
    PoolAlloc p;
    p.init(64 * U);
    std::mt19937 rng(7);
    std::map<size_t, size_t> live;  // off -> units
    for (int it = 0; it < 20000; ++it) {
        if (live.empty() || (rng() % 3 != 0 && live.size() < 40)) {
            const size_t units = 1 + rng() % 6;
            size_t off = 0;
            if (p.alloc(units * U - (rng() % U), &off)) {  // sizes round up to 256
                CHECK(off % U == 0 && off + units * U <= 64 * U);
                for (auto& kv : live) CHECK(off + units * U <= kv.first || kv.first + kv.second * U <= off);
                live[off] = units;
            }
        } else {
            auto it2 = live.begin();
            std::advance(it2, rng() % live.size());
            p.release(it2->first, it2->second * U);
            live.erase(it2);
        }
        size_t used = 0;
        for (auto& kv : live) used += kv.second * U;
        CHECK(p.free_bytes() + used == 64 * U);
    }
    for (auto& kv : live) p.release(kv.first, kv.second * U);
    size_t off = 0;
    CHECK(p.alloc(64 * U, &off) && off == 0);  // fully coalesced again
    std::puts("ok pool_alloc");
}

static void admission_and_refusal() {
    FakeClock clk;
    Core c(cfg(10, 10), clk.fn());
    uint64_t a = 0, b = 0, d = 0, big = 0;
    CHECK(c.submit(job(11), &big) == RT_ERR_BUSY);  // never fits the device pool
    CHECK(c.submit(job(1, 11), &big) == RT_ERR_BUSY);  // never fits the staging pool
    CHECK(c.submit(job(4), &a) == RT_OK && c.submit(job(4), &b) == RT_OK && c.submit(job(4), &d) == RT_OK);
    auto ad = admit(c);
    CHECK(ad.size() == 2 && ad[0].id == a && ad[1].id == b);
    CHECK(ad[0].dev_off != ad[1].dev_off);
    CHECK(state(c, a) == State::Running && state(c, d) == State::Pending);
    CHECK(admit(c).empty());
    c.finished(a, RT_OK);
    CHECK(state(c, a) == State::Done);
    ad = admit(c);
    CHECK(ad.size() == 1 && ad[0].id == d);
    int done = 0;
    CHECK(c.poll(a, &done) == RT_OK && done == 1);
    CHECK(c.poll(d, &done) == RT_OK && done == 0);
    CHECK(c.poll(999, &done) == RT_ERR_NOT_FOUND);
    std::puts("ok admission_and_refusal");
}

static void overtaking_and_age_limit() {
    FakeClock clk;
    Core c(cfg(10, 10), clk.fn());
    uint64_t r1, r2, big, s1, s2;
    c.submit(job(5), &r1);
    c.submit(job(4), &r2);
    CHECK(admit(c).size() == 2);  // 1 unit left
    c.submit(job(6), &big);       // does not fit yet
    c.submit(job(1), &s1);        // fits: passes the young big request
    auto ad = admit(c);
    CHECK(ad.size() == 1 && ad[0].id == s1);
    clk.t += 25000;  // big has now waited longer than starve_us
    c.finished(s1, RT_OK);
    c.submit(job(1), &s2);
    CHECK(admit(c).empty());  // nothing passes the old request any more
    c.finished(r1, RT_OK);
    CHECK(admit(c).empty());  // 6 units free, but not in one piece
    c.finished(r2, RT_OK);
    ad = admit(c);
    CHECK(ad.size() == 2 && ad[0].id == big && ad[1].id == s2);
    std::puts("ok overtaking_and_age_limit");
}

static void max_inflight() {
    FakeClock clk;
    Core c(cfg(100, 100, 2), clk.fn());
    uint64_t ids[4];
    for (auto& i : ids) c.submit(job(1), &i);
    CHECK(admit(c).size() == 2);
    c.finished(ids[0], RT_OK);
    auto ad = admit(c);
    CHECK(ad.size() == 1 && ad[0].id == ids[2]);
    std::puts("ok max_inflight");
}

static void cancel_in_every_state() {
    FakeClock clk;
    Core c(cfg(10, 10), clk.fn());
    uint64_t p, r, d;
    c.submit(job(4, 4), &r);
    c.submit(job(4, 4), &d);
    CHECK(admit(c).size() == 2);
    c.submit(job(4, 4), &p);  // pending: no room
    // pending: removed at once, never admitted
    CHECK(c.cancel(p) == RT_OK && state(c, p) == State::Cancelled);
    CHECK(c.wait(p, 0) == RT_ERR_CANCELLED);
    // running: reported at once, blocks held until the device finishes
    const Counts before = c.counts();
    CHECK(c.cancel(r) == RT_OK && state(c, r) == State::Cancelled);
    CHECK(c.wait(r, 0) == RT_ERR_CANCELLED);
    CHECK(c.counts().dev_free == before.dev_free);
    size_t off = 0;
    CHECK(c.take_begin(r, &off) == RT_ERR_CANCELLED);  // forgotten once the device lets go
    c.finished(r, RT_OK);
    CHECK(!c.find_state(r, nullptr));
    CHECK(c.counts().dev_free == before.dev_free + 4 * U && c.counts().host_free == before.host_free + 4 * U);
    // done: results dropped
    c.finished(d, RT_OK);
    const size_t host_free = c.counts().host_free;
    CHECK(c.cancel(d) == RT_OK && state(c, d) == State::Cancelled);
    CHECK(c.counts().host_free == host_free + 4 * U);
    CHECK(c.take_begin(d, &off) == RT_ERR_CANCELLED && !c.find_state(d, nullptr));
    CHECK(c.cancel(d) == RT_ERR_NOT_FOUND);
    CHECK(c.counts().dev_free == 10 * U && c.counts().host_free == 10 * U);
    CHECK(admit(c).empty());  // the cancelled pending request never runs
    std::puts("ok cancel_in_every_state");
}

static void deadlines() {
    FakeClock clk;
    Core c(cfg(10, 10), clk.fn());
    const int64_t t0 = clk.t;
    uint64_t run, pend, late;
    c.submit(job(6, 2, false, t0 + 5000), &run);
    CHECK(admit(c).size() == 1);
    c.submit(job(6, 2, false, t0 + 3000), &pend);  // cannot fit while run holds 6
    c.submit(job(1, 1, false, t0 + 100000), &late);
    std::vector<uint64_t> gone;
    clk.t = t0 + 2999;
    c.expire(&gone);
    CHECK(gone.empty());
    clk.t = t0 + 3000;
    c.expire(&gone);
    CHECK(gone.size() == 1 && gone[0] == pend && state(c, pend) == State::Expired);
    CHECK(c.wait(pend, 0) == RT_ERR_EXPIRED);
    clk.t = t0 + 6000;
    gone.clear();
    c.expire(&gone);
    CHECK(gone.size() == 1 && gone[0] == run && state(c, run) == State::Expired);  // running: at once
    const size_t dev_free = c.counts().dev_free;
    c.finished(run, RT_OK);  // its results are dropped, its blocks freed
    CHECK(state(c, run) == State::Expired && c.counts().dev_free == dev_free + 6 * U);
    auto ad = admit(c);
    CHECK(ad.size() == 1 && ad[0].id == late);
    c.finished(late, RT_OK);
    clk.t = t0 + 200000;  // a deadline after completion changes nothing
    c.expire(&gone);
    CHECK(state(c, late) == State::Done);
    size_t off;
    CHECK(c.take_begin(run, &off) == RT_ERR_EXPIRED);
    CHECK(c.take_begin(pend, &off) == RT_ERR_EXPIRED);
    CHECK(c.take_begin(late, &off) == RT_OK);
    c.take_end(late);
    CHECK(c.counts().dev_free == 10 * U && c.counts().host_free == 10 * U);
    std::puts("ok deadlines");
}

static void lanes() {
    FakeClock clk;
    Config k = cfg(1000, 1000);
    k.lanes = 4;
    k.kerr_only_lanes = 1;
    Core c(k, clk.fn());
    uint64_t m1, m2, m3, m4, k1, k2;
    c.submit(job(1, 1, true), &m1);
    c.submit(job(1, 1, true), &m2);
    c.submit(job(1, 1, true), &m3);
    c.submit(job(1, 1, true), &m4);  // three shared lanes: the fourth superposition waits
    c.submit(job(1, 1, false), &k1);  // the single-Kerr lane is still free
    c.submit(job(1, 1, false), &k2);  // no lane left
    auto ad = admit(c);
    std::map<uint64_t, uint32_t> lane;
    for (auto& x : ad) lane[x.id] = x.lane;
    CHECK(ad.size() == 4 && lane.count(m1) && lane.count(m2) && lane.count(m3) && lane.count(k1));
    CHECK(lane[k1] == 3 && lane[m1] != 3 && lane[m2] != 3 && lane[m3] != 3);
    CHECK(state(c, m4) == State::Pending && state(c, k2) == State::Pending && c.lanes_busy() == 4);
    c.finished(k1, RT_OK);  // frees the single-Kerr lane: k2 takes it, m4 cannot
    ad = admit(c);
    CHECK(ad.size() == 1 && ad[0].id == k2 && ad[0].lane == 3);
    c.finished(m2, RT_OK);
    ad = admit(c);
    CHECK(ad.size() == 1 && ad[0].id == m4 && ad[0].lane == lane[m2]);
    std::puts("ok lanes");
}

static void take_protocol() {
    FakeClock clk;
    Core c(cfg(10, 10), clk.fn());
    uint64_t r;
    c.submit(job(2, 3), &r);
    size_t off = 99;
    CHECK(c.take_begin(r, &off) == RT_ERR_BUSY);  // pending
    auto ad = admit(c);
    CHECK(c.take_begin(r, &off) == RT_ERR_BUSY);  // running
    c.finished(r, RT_OK);
    CHECK(c.take_begin(r, &off) == RT_OK && off == ad[0].host_off);
    CHECK(c.take_begin(r, &off) == RT_ERR_BUSY);  // a second taker waits its turn
    CHECK(c.cancel(r) == RT_OK && state(c, r) == State::Done);  // the take in progress wins
    c.take_end(r, /*give_back=*/true);
    CHECK(state(c, r) == State::Done && c.counts().host_free == 7 * U);
    CHECK(c.take_begin(r, &off) == RT_OK);
    c.take_end(r);
    CHECK(!c.find_state(r, nullptr) && c.counts().host_free == 10 * U);
    CHECK(c.take_begin(r, &off) == RT_ERR_NOT_FOUND);
    // a device error
    uint64_t f;
    c.submit(job(1), &f);
    admit(c);
    c.finished(f, RT_ERR_DEVICE);
    CHECK(state(c, f) == State::Failed && c.wait(f, 0) == RT_ERR_DEVICE);
    CHECK(c.take_begin(f, &off) == RT_ERR_DEVICE && !c.find_state(f, nullptr));
    uint64_t g;
    c.submit(job(1), &g);
    admit(c);
    c.launch_failed(g, RT_ERR_OOM);
    CHECK(state(c, g) == State::Failed && c.counts().dev_free == 10 * U && c.counts().host_free == 10 * U);
    std::puts("ok take_protocol");
}

static void wait_times_out_and_wakes() {
    FakeClock clk;
    Core c(cfg(10, 10), clk.fn());
    uint64_t r;
    c.submit(job(1), &r);
    CHECK(c.wait(r, 10) == RT_ERR_TIMEOUT);
    admit(c);
    std::thread dev([&] {
        std::this_thread::sleep_for(std::chrono::milliseconds(20));
        c.finished(r, RT_OK);
    });
    CHECK(c.wait(r, 5000) == RT_OK);
    dev.join();
    std::puts("ok wait_times_out_and_wakes");
}

static void shutdown_wakes_waiters() {
    FakeClock clk;
    Core c(cfg(2, 2), clk.fn());
    uint64_t r, p;
    c.submit(job(2, 2), &r);
    admit(c);
    c.submit(job(2, 2), &p);
    std::atomic<int> st{-1};
    std::thread w([&] { st = c.wait(p, -1); });
    std::this_thread::sleep_for(std::chrono::milliseconds(10));
    c.shutdown();
    w.join();
    CHECK(st == RT_ERR_CANCELLED);
    uint64_t q;
    CHECK(c.submit(job(1), &q) == RT_ERR_CANCELLED);
    CHECK(admit(c).empty());
    c.finished(r, RT_OK);  // the running one still completes
    CHECK(state(c, r) == State::Done);
    std::puts("ok shutdown_wakes_waiters");
}

// Wave slots: grants never exceed the free slots; a single-Kerr request takes
// the free slots up to its need; a superposition at most three quarters of
// all slots and never the last quarter; a request without room waits while
// later ones that fit pass it; slots come back when the kernel ends; a cancel
// asks the running kernel to stop (Resize to 0) but keeps its slots until it
// ends.
static Job mjob(uint32_t need) {
    Job j = sjob(need);
    j.multi = true;
    return j;
}

static void wave_slots() {
    FakeClock clk;
    Config k = cfg(1000, 1000);
    k.slots = 100;
    k.lanes = 5;
    k.kerr_only_lanes = 1;
    Core c(k, clk.fn());
    uint64_t m1, m2, m3, k1, k2;
    std::vector<Resize> rs;
    std::vector<Admission> ad;
    auto adm = [&]() {
        ad.clear();
        rs.clear();
        c.admit(&ad, &rs);
    };
    c.submit(mjob(500), &m1);
    c.submit(mjob(500), &m2);
    c.submit(mjob(500), &m3);
    c.submit(sjob(30), &k1);
    c.submit(sjob(1), &k2);
    adm();
    // m1: three quarters of all slots; m2, m3: nothing free above the quarter
    // kept back; k1: the free slots (the kept quarter); k2: none left
    CHECK(ad.size() == 2 && ad[0].id == m1 && ad[0].slots == 75 && ad[1].id == k1 && ad[1].slots == 25);
    CHECK(c.counts().slots_free == 0);
    CHECK(state(c, m2) == State::Pending && state(c, m3) == State::Pending && state(c, k2) == State::Pending);
    CHECK(rs.empty());
    c.finished(k1, RT_OK);
    adm();
    CHECK(ad.size() == 1 && ad[0].id == k2 && ad[0].slots == 1);  // m2, m3 still find no room
    c.finished(m1, RT_OK);
    adm();
    CHECK(ad.size() == 1 && ad[0].id == m2 && ad[0].slots == 74);  // 99 free: 74 above the quarter
    CHECK(state(c, m3) == State::Pending);
    c.cancel(m2);
    adm();
    CHECK(rs.size() == 1 && rs[0].id == m2 && rs[0].slots == 0);
    CHECK(c.counts().slots_free == 100 - 74 - 1);  // its slots stay until its kernel ends
    CHECK(ad.empty() && state(c, m3) == State::Pending);
    c.finished(m2, RT_OK);
    adm();
    CHECK(ad.size() == 1 && ad[0].id == m3 && ad[0].slots == 74);
    c.finished(m3, RT_OK);
    c.finished(k2, RT_OK);
    CHECK(c.counts().slots_free == 100 && c.lanes_busy() == 0);
    uint64_t bad;
    c.submit(sjob(1000), &bad);
    adm();
    CHECK(ad.size() == 1 && ad[0].slots == 100);
    c.launch_failed(bad, RT_ERR_DEVICE);
    CHECK(c.counts().slots_free == 100);
    {
        // a grant below min(need, slots / 8) is not worth starting: the request waits
        Config k3 = cfg(1000, 1000);
        k3.slots = 100;
        k3.lanes = 4;
        Core c3(k3, clk.fn());
        uint64_t a1, a2, a3;
        c3.submit(sjob(95), &a1);
        c3.submit(sjob(40), &a2);  // 5 free < 12: waits
        c3.submit(sjob(3), &a3);   // needs only 3: starts
        std::vector<Admission> a;
        c3.admit(&a, nullptr);
        CHECK(a.size() == 2 && a[0].id == a1 && a[1].id == a3 && a[1].slots == 3);
        c3.finished(a1, RT_OK);
        a.clear();
        c3.admit(&a, nullptr);
        CHECK(a.size() == 1 && a[0].id == a2 && a[0].slots == 40);
    }
    std::puts("ok wave_slots");
}

// Clients submit, cancel, wait and take; a fake device admits and finishes
// requests after short delays; a clock thread moves time on. Every request
// must end exactly once and every block must come back.
static void concurrent_run() {
    FakeClock clk;
    Config kc = cfg(64, 48, 16);
    kc.slots = 64;
    kc.lanes = 4;
    kc.kerr_only_lanes = 1;
    Core c(kc, clk.fn());
    std::atomic<bool> stop{false};
    std::atomic<int> device_done{0};
    std::thread ticker([&] {
        while (!stop) {
            clk.t += 500;
            std::this_thread::sleep_for(std::chrono::microseconds(200));
        }
    });
    std::thread device([&] {
        std::mt19937 rng(11);
        std::vector<std::pair<uint64_t, int64_t>> live;  // id, finish time
        while (!stop || !live.empty()) {
            c.expire();
            std::vector<Admission> ad;
            c.admit(&ad);
            for (auto& a : ad) {
                if (rng() % 50 == 0) {
                    c.launch_failed(a.id, RT_ERR_DEVICE);
                    continue;
                }
                live.emplace_back(a.id, clk.t + 200 + static_cast<int64_t>(rng() % 3000));
            }
            for (size_t i = 0; i < live.size();) {
                if (clk.t >= live[i].second) {
                    c.finished(live[i].first, rng() % 100 == 0 ? RT_ERR_DEVICE : RT_OK);
                    ++device_done;
                    live[i] = live.back();
                    live.pop_back();
                } else {
                    ++i;
                }
            }
            c.wait_work(100);
        }
    });
    const int kClients = 6, kPer = 300;
    std::atomic<int> outcomes[10] = {};
    std::vector<std::thread> clients;
    for (int t = 0; t < kClients; ++t) {
        clients.emplace_back([&, t] {
            std::mt19937 rng(100 + t);
            for (int k = 0; k < kPer; ++k) {
                const bool multi = rng() % 4 == 0;
                const int64_t deadline = (rng() % 5 == 0) ? clk.t + 500 + rng() % 4000 : 0;
                uint64_t id = 0;
                Job jb = job(1 + rng() % 12, 1 + rng() % 8, multi, deadline);
                jb.slots_need = 1 + rng() % 80;
                const rt_status s = c.submit(jb, &id);
                if (s != RT_OK) {
                    ++outcomes[s];
                    continue;
                }
                if (rng() % 7 == 0) c.cancel(id);
                rt_status w = c.wait(id, 20000);
                if (w == RT_ERR_TIMEOUT) {  // should not happen: the device always makes progress
                    CHECK(false);
                    c.cancel(id);
                    w = c.wait(id, -1);
                }
                if (rng() % 9 == 0) c.cancel(id);  // after the fact
                size_t off = 0;
                rt_status tk;
                while ((tk = c.take_begin(id, &off)) == RT_ERR_BUSY) std::this_thread::yield();
                if (tk == RT_OK) c.take_end(id);
                ++outcomes[tk];
            }
        });
    }
    for (auto& th : clients) th.join();
    stop = true;
    c.notify_worker();
    device.join();
    ticker.join();
    int total = 0;
    for (auto& o : outcomes) total += o;
    CHECK(total == kClients * kPer);
    const Counts k = c.counts();
    CHECK(k.pending == 0 && k.running == 0 && k.finished == 0);
    CHECK(k.dev_free == 64 * U && k.host_free == 48 * U && k.slots_free == 64);
    std::printf("ok concurrent_run (%d requests: done %d, cancelled %d, expired %d, failed %d; device finished %d)\n",
                total, outcomes[RT_OK].load(), outcomes[RT_ERR_CANCELLED].load(), outcomes[RT_ERR_EXPIRED].load(),
                outcomes[RT_ERR_DEVICE].load(), device_done.load());
}

int main() {
    pool_alloc();
    admission_and_refusal();
    overtaking_and_age_limit();
    max_inflight();
    cancel_in_every_state();
    deadlines();
    lanes();
    take_protocol();
    wait_times_out_and_wakes();
    shutdown_wakes_waiters();
    wave_slots();
    concurrent_run();
    if (g_fail) {
        std::fprintf(stderr, "%d check(s) failed\n", g_fail.load());
        return 1;
    }
    std::puts("all serve core checks passed");
    return 0;
}
