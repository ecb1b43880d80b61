// rt_serve_core.cpp — request states, admission and memory accounting of the
// serving loop (no device calls; see rt_serve_core.hpp).
#include "rt_serve_core.hpp"

#include <algorithm>
#include <chrono>
#include <iterator>
#include <limits>

namespace rt {
namespace serve {

namespace {
size_t al256(size_t v) { return (v + 255) & ~size_t(255); }
}  // namespace

int64_t steady_now_us() {
    return std::chrono::duration_cast<std::chrono::microseconds>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

// ------------------------------------------------------------ PoolAlloc
void PoolAlloc::init(size_t size) {
    size_ = size & ~size_t(255);
    free_.clear();
    if (size_) free_[0] = size_;
}

bool PoolAlloc::alloc(size_t bytes, size_t* off) {
    bytes = al256(bytes == 0 ? 1 : bytes);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
        if (it->second >= bytes) {
            *off = it->first;
            const size_t rest = it->second - bytes;
            const size_t noff = it->first + bytes;
            free_.erase(it);
            if (rest) free_[noff] = rest;
            return true;
        }
    }
    return false;
}

void PoolAlloc::release(size_t off, size_t bytes) {
    bytes = al256(bytes == 0 ? 1 : bytes);
    auto it = free_.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
        it->second += nx->second;
        free_.erase(nx);
    }
    if (it != free_.begin()) {
        auto pv = std::prev(it);
        if (pv->first + pv->second == it->first) {
            pv->second += it->second;
            free_.erase(it);
        }
    }
}

size_t PoolAlloc::free_bytes() const {
    size_t s = 0;
    for (const auto& kv : free_) s += kv.second;
    return s;
}

// ----------------------------------------------------------------- Core
Core::Core(const Config& cfg, Clock clock) : cfg_(cfg), clock_(std::move(clock)) {
    if (cfg_.lanes == 0) cfg_.lanes = 1;
    if (cfg_.kerr_only_lanes >= cfg_.lanes) cfg_.kerr_only_lanes = cfg_.lanes - 1;
    dev_.init(cfg_.dev_bytes);
    host_.init(cfg_.host_bytes);
    lane_busy_.assign(cfg_.lanes, false);
}

rt_status Core::submit(const Job& job, uint64_t* id) {
    if (!id) return RT_ERR_INVALID_ARG;
    if (al256(job.dev_bytes) > dev_.size() || al256(job.host_bytes) > host_.size()) return RT_ERR_BUSY;
    std::lock_guard<std::mutex> lk(mu_);
    if (stopped_) return RT_ERR_CANCELLED;
    const uint64_t rid = next_id_++;
    Rec& r = recs_[rid];
    r.job = job;
    r.submit_us = clock_();
    pending_.push_back(rid);
    *id = rid;
    work_flag_ = true;
    cv_work_.notify_all();
    return RT_OK;
}

void Core::release_host(Rec& r) {
    if (r.has_host) {
        host_.release(r.host_off, r.job.host_bytes);
        r.has_host = false;
        work_flag_ = true;
        cv_work_.notify_all();
    }
}

void Core::end_locked(uint64_t id, Rec& r, State s, rt_status err) {
    if (r.state == State::Pending) pending_.erase(std::remove(pending_.begin(), pending_.end(), id), pending_.end());
    if (r.state == State::Running && s != State::Done) {
        // its results will be dropped: its waves may all leave now (its slots
        // come back when its kernel ends)
        resizes_.push_back(Resize{id, 0});
        work_flag_ = true;
        cv_work_.notify_all();
    }
    r.state = s;
    r.err = err;
    if (s != State::Done && !r.on_device) release_host(r);
    cv_done_.notify_all();
}

void Core::maybe_erase(std::map<uint64_t, Rec>::iterator it) {
    Rec& r = it->second;
    if (r.forget && !r.on_device && !r.taking) {
        release_host(r);
        recs_.erase(it);
    }
}

rt_status Core::cancel(uint64_t id) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = recs_.find(id);
    if (it == recs_.end()) return RT_ERR_NOT_FOUND;
    Rec& r = it->second;
    if (r.taking) return RT_OK;  // the results are being handed over already
    if (r.state == State::Pending || r.state == State::Running || r.state == State::Done)
        end_locked(id, r, State::Cancelled, RT_ERR_CANCELLED);
    work_flag_ = true;
    cv_work_.notify_all();
    return RT_OK;
}

rt_status Core::poll(uint64_t id, int* done) {
    if (!done) return RT_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(mu_);
    auto it = recs_.find(id);
    if (it == recs_.end() || it->second.forget) return RT_ERR_NOT_FOUND;
    *done = finished_state(it->second.state) ? 1 : 0;
    return RT_OK;
}

// Timed waits run on the system clock (wait_until): the steady-clock timed wait
// of libstdc++ goes through pthread_cond_clockwait, which the ThreadSanitizer
// of this toolchain does not intercept, so the sanitizer runs would not see
// the mutex released inside the wait. A wall-clock jump can only lengthen or
// shorten one wait.
rt_status Core::wait(uint64_t id, int64_t timeout_ms) {
    std::unique_lock<std::mutex> lk(mu_);
    auto over = [&]() {
        auto it = recs_.find(id);
        return it == recs_.end() || it->second.forget || finished_state(it->second.state);
    };
    if (timeout_ms < 0)
        cv_done_.wait(lk, over);
    else if (!cv_done_.wait_until(lk, std::chrono::system_clock::now() + std::chrono::milliseconds(timeout_ms), over))
        return RT_ERR_TIMEOUT;
    auto it = recs_.find(id);
    if (it == recs_.end() || it->second.forget) return RT_ERR_NOT_FOUND;
    return it->second.err;
}

rt_status Core::take_begin(uint64_t id, size_t* host_off) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = recs_.find(id);
    if (it == recs_.end() || it->second.forget) return RT_ERR_NOT_FOUND;
    Rec& r = it->second;
    if (!finished_state(r.state)) return RT_ERR_BUSY;
    if (r.taking) return RT_ERR_BUSY;
    if (r.state != State::Done) {
        const rt_status e = r.err;
        r.forget = true;
        maybe_erase(it);
        return e;
    }
    r.taking = true;
    if (host_off) *host_off = r.host_off;
    return RT_OK;
}

void Core::take_end(uint64_t id, bool give_back) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = recs_.find(id);
    if (it == recs_.end()) return;
    Rec& r = it->second;
    r.taking = false;
    if (give_back) return;
    r.forget = true;
    maybe_erase(it);
}

void Core::expire(std::vector<uint64_t>* dropped) {
    std::lock_guard<std::mutex> lk(mu_);
    const int64_t now = clock_();
    for (auto& kv : recs_) {
        Rec& r = kv.second;
        if ((r.state == State::Pending || r.state == State::Running) && r.job.deadline_us != 0 &&
            now >= r.job.deadline_us) {
            end_locked(kv.first, r, State::Expired, RT_ERR_EXPIRED);
            if (dropped) dropped->push_back(kv.first);
        }
    }
}

uint32_t Core::take_slots(uint32_t need, bool multi) {
    const uint32_t free = cfg_.slots - slots_used_;
    uint32_t avail = free;
    if (multi) {
        const uint32_t share = cfg_.multi_share ? cfg_.multi_share : std::max<uint32_t>(1, cfg_.slots * 3 / 4);
        const uint32_t reserve = cfg_.multi_reserve ? cfg_.multi_reserve : cfg_.slots / 4;
        avail = (free > reserve) ? std::min(free - reserve, share) : 0;
    }
    const uint32_t floor = std::min(need, cfg_.min_grant ? cfg_.min_grant : std::max<uint32_t>(1, cfg_.slots / 8));
    const uint32_t grant = std::min(need, avail);
    if (grant < floor) return 0;
    slots_used_ += grant;
    return grant;
}

void Core::admit(std::vector<Admission>* out, std::vector<Resize>* resized) {
    std::lock_guard<std::mutex> lk(mu_);
    work_flag_ = false;
    if (resized) resized->insert(resized->end(), resizes_.begin(), resizes_.end());
    resizes_.clear();
    if (stopped_) return;
    const int64_t now = clock_();
    for (auto pit = pending_.begin(); pit != pending_.end();) {
        if (running_ >= cfg_.max_inflight) break;
        const uint64_t id = *pit;
        Rec& r = recs_.at(id);
        // a free lane: single-Kerr requests try their own lanes first
        uint32_t lane = cfg_.lanes;
        const uint32_t shared = cfg_.lanes - cfg_.kerr_only_lanes;
        if (!r.job.multi)
            for (uint32_t l = shared; l < cfg_.lanes && lane == cfg_.lanes; ++l)
                if (!lane_busy_[l]) lane = l;
        for (uint32_t l = 0; l < shared && lane == cfg_.lanes; ++l)
            if (!lane_busy_[l]) lane = l;
        if (lane == cfg_.lanes) {  // every lane it may use is busy
            ++pit;
            continue;
        }
        size_t doff = 0, hoff = 0;
        const bool dev_ok = dev_.alloc(r.job.dev_bytes, &doff);
        const bool host_ok = dev_ok && host_.alloc(r.job.host_bytes, &hoff);
        if (!host_ok) {
            if (dev_ok) dev_.release(doff, r.job.dev_bytes);
            // an old request that does not fit holds back everything behind it
            if (now - r.submit_us >= cfg_.starve_us) break;
            ++pit;
            continue;
        }
        // wave slots: the free ones up to the need (superpositions: capped)
        uint32_t grant = 0;
        if (cfg_.slots != 0) {
            grant = take_slots(std::max<uint32_t>(r.job.slots_need, 1), r.job.multi);
            if (grant == 0) {  // no room for it now; a later single-Kerr request may still fit
                dev_.release(doff, r.job.dev_bytes);
                host_.release(hoff, r.job.host_bytes);
                ++pit;
                continue;
            }
        }
        r.state = State::Running;
        r.on_device = true;
        r.has_host = true;
        r.dev_off = doff;
        r.host_off = hoff;
        lane_busy_[lane] = true;
        r.lane = lane;
        r.slots = grant;
        ++running_;
        pit = pending_.erase(pit);
        out->push_back(Admission{id, doff, hoff, lane, grant});
    }
}

void Core::launch_failed(uint64_t id, rt_status st) { finished(id, st == RT_OK ? RT_ERR_DEVICE : st); }

void Core::finished(uint64_t id, rt_status st) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = recs_.find(id);
    if (it == recs_.end()) return;
    Rec& r = it->second;
    if (!r.on_device) return;
    dev_.release(r.dev_off, r.job.dev_bytes);
    r.on_device = false;
    lane_busy_[r.lane] = false;
    slots_used_ -= r.slots;
    r.slots = 0;
    --running_;
    if (r.state == State::Running) {
        r.state = (st == RT_OK) ? State::Done : State::Failed;
        r.err = st;
        if (st != RT_OK) release_host(r);
    } else {
        // cancelled or expired while on the device: drop the results
        release_host(r);
    }
    work_flag_ = true;
    cv_work_.notify_all();
    cv_done_.notify_all();
    maybe_erase(it);
}

void Core::wait_work(int64_t timeout_us) {
    std::unique_lock<std::mutex> lk(mu_);
    cv_work_.wait_until(lk, std::chrono::system_clock::now() + std::chrono::microseconds(timeout_us),
                        [&] { return work_flag_ || stopped_; });
}

void Core::notify_worker() {
    std::lock_guard<std::mutex> lk(mu_);
    work_flag_ = true;
    cv_work_.notify_all();
}

void Core::shutdown() {
    std::lock_guard<std::mutex> lk(mu_);
    stopped_ = true;
    while (!pending_.empty()) {
        const uint64_t id = pending_.front();
        end_locked(id, recs_.at(id), State::Cancelled, RT_ERR_CANCELLED);
    }
    cv_work_.notify_all();
    cv_done_.notify_all();
}

Counts Core::counts() const {
    std::lock_guard<std::mutex> lk(mu_);
    Counts c;
    for (const auto& kv : recs_) {
        const Rec& r = kv.second;
        if (r.forget) continue;
        if (r.state == State::Pending)
            ++c.pending;
        else if (r.state == State::Running)
            ++c.running;
        else
            ++c.finished;
    }
    c.dev_free = dev_.free_bytes();
    c.host_free = host_.free_bytes();
    c.slots_free = cfg_.slots - slots_used_;
    return c;
}

bool Core::find_state(uint64_t id, State* st) const {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = recs_.find(id);
    if (it == recs_.end() || it->second.forget) return false;
    if (st) *st = it->second.state;
    return true;
}

uint32_t Core::lanes_busy() const {
    std::lock_guard<std::mutex> lk(mu_);
    uint32_t n = 0;
    for (bool b : lane_busy_) n += b ? 1u : 0u;
    return n;
}

}  // namespace serve
}  // namespace rt
