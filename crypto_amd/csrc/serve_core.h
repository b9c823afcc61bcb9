// The scheduler of the commitment serving loop (include/kinet/crypto/kzg_serve.h),
// independent of the device: submitters append tickets to a FIFO; one
// scheduler thread owns the executor's lanes and loops
//   1. completions: a busy lane whose batch finished hands its results to the
//      tickets and returns its footprint to the HBM budget;
//   2. dispatch: while a lane is free, it takes the oldest pending ticket and,
//      in arrival order, every further pending ticket of the same curve that
//      fits the batch limits (continuous batching: whatever is waiting when a
//      lane frees up rides together); the batch is admitted only if its
//      device footprint fits the budget left by the batches in flight (a
//      batch is always admitted when nothing is in flight), else it waits for
//      a completion (admission control).
// A batch whose error flag is set (some evaluation >= r) is re-run one
// request per batch, so only the offending request fails; a batch whose
// launch throws (staging allocation, launch failure) or whose device work
// fails fails its own requests only.
//
// Exec supplies the device side:
//   size_t lanes() const
//   uint64_t footprint(int curve, const std::vector<uint32_t>& sizes)
//   void launch(size_t lane, int curve, const std::vector<std::pair<const uint8_t*, uint32_t>>& src,
//               const std::vector<uint32_t>& sizes)        // may throw
//   int poll(size_t lane)      // 0 running, 1 finished, < 0 device failure
//   bool error_flag(size_t lane, size_t count)             // after poll() == 1
//   const uint8_t* result(size_t lane, size_t k)           // compressed commitment k
// The HIP library instantiates it with its CommitLane executor (serve.hip);
// the host check instantiates it with a fake one that injects failures.
#pragma once

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <utility>
#include <vector>

#include "kinet/crypto/kzg_serve.h"

namespace kz {

// crypto status codes (c-abi/crypto.h:60-66) used by the scheduler
constexpr int kServeErrInput = -1;
constexpr int kServeErrBackend = -4;

template <class Exec>
class ServeCore {
public:
    // max_n[curve]: the largest request of a served curve (its SRS size), 0 = not served
    ServeCore(Exec& exec, const kzg_server_config& cfg, const size_t (&max_n)[3]) : exec_(exec), cfg_(cfg) {
        for (int c = 0; c < 3; ++c) max_n_[c] = max_n[c];
        lanes_.resize(exec_.lanes());
        sched_ = std::thread([this] { run(); });
    }
    ~ServeCore() {
        {
            std::lock_guard<std::mutex> g(mu_);
            stop_ = true;
        }
        cv_sched_.notify_all();
        if (sched_.joinable()) sched_.join();
    }
    ServeCore(const ServeCore&) = delete;
    ServeCore& operator=(const ServeCore&) = delete;

    // Returns a ticket >= 0, or kServeErrInput.
    int64_t submit(uint32_t curve, const uint8_t* evals, size_t n) {
        if (!evals || (curve != 1 && curve != 2) || !max_n_[curve]) return kServeErrInput;
        if (n == 0 || (n & (n - 1)) != 0 || n > max_n_[curve] || n > cfg_.max_batch_elems || n > (1u << 26))
            return kServeErrInput;
        int64_t tk;
        {
            std::lock_guard<std::mutex> g(mu_);
            if (stop_) return kServeErrInput;
            tk = next_ticket_++;
            Request& r = reqs_[tk];
            r.curve = (int)curve;
            r.evals = evals;
            r.n = (uint32_t)n;
            r.t_submit = now_ns();
            pending_.push_back(tk);
            ++stats_.submitted;
        }
        cv_sched_.notify_one();
        return tk;
    }

    // Blocks until the ticket completes; 0 with the commitment, or its error.
    int wait(int64_t ticket, uint8_t* commit_out, uint64_t* t_submit_ns, uint64_t* t_done_ns) {
        std::unique_lock<std::mutex> lk(mu_);
        auto it = reqs_.find(ticket);
        if (it == reqs_.end()) return kServeErrInput;
        cv_done_.wait(lk, [&] { return it->second.status != 0; });  // std::map: `it` survives inserts
        const Request r = it->second;
        reqs_.erase(it);
        lk.unlock();
        if (t_submit_ns) *t_submit_ns = r.t_submit;
        if (t_done_ns) *t_done_ns = r.t_done;
        if (r.status < 0) return r.status;
        std::memcpy(commit_out, r.commit, commit_bytes(r.curve));
        return 0;
    }

    kzg_server_stats stats() {
        std::lock_guard<std::mutex> g(mu_);
        return stats_;
    }

    // Compressed size of a curve's commitment.
    static size_t commit_bytes(int curve) { return curve == 2 ? 48 : 32; }

private:
    struct Request {
        int curve = 0;
        const uint8_t* evals = nullptr;
        uint32_t n = 0;
        bool solo = false;  // run alone (re-run after a failed batch)
        int status = 0;     // 0 pending, 1 done, < 0 error
        uint64_t t_submit = 0, t_done = 0;
        uint8_t commit[48];
    };
    struct Lane {
        bool busy = false;
        int curve = 0;
        uint64_t bytes = 0;
        std::vector<int64_t> tickets;  // in batch (size-sorted) order
    };

    uint64_t now_ns() const {
        return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now() - t0_)
            .count();
    }

    void fail(Lane& lane, int status) {
        const uint64_t t = now_ns();
        for (int64_t tk : lane.tickets) {
            Request& r = reqs_[tk];
            r.status = status;
            r.t_done = t;
            ++stats_.failed;
        }
    }

    void release(Lane& lane) {
        inflight_ -= lane.bytes;
        lane.busy = false;
        lane.tickets.clear();
    }

    // under mu_
    bool finish_lanes() {
        bool any = false;
        for (size_t li = 0; li < lanes_.size(); ++li) {
            Lane& lane = lanes_[li];
            if (!lane.busy) continue;
            const int q = exec_.poll(li);
            if (q == 0) continue;
            any = true;
            const bool err = q > 0 && exec_.error_flag(li, lane.tickets.size());
            if (q < 0) {
                fail(lane, kServeErrBackend);
            } else if (err && lane.tickets.size() == 1) {
                fail(lane, kServeErrInput);
            } else if (err) {
                // some evaluation >= r: re-run each request alone, in order, at the front
                for (auto it = lane.tickets.rbegin(); it != lane.tickets.rend(); ++it) {
                    reqs_[*it].solo = true;
                    pending_.push_front(*it);
                }
                ++stats_.rerun_batches;
            } else {
                const uint64_t t = now_ns();
                const size_t cb = commit_bytes(lane.curve);
                for (size_t k = 0; k < lane.tickets.size(); ++k) {
                    Request& r = reqs_[lane.tickets[k]];
                    std::memcpy(r.commit, exec_.result(li, k), cb);
                    r.status = 1;
                    r.t_done = t;
                    ++stats_.completed;
                }
            }
            release(lane);
        }
        if (any) cv_done_.notify_all();
        return any;
    }

    // under mu_ (released while staging and launching)
    bool dispatch(std::unique_lock<std::mutex>& lk) {
        bool any = false;
        for (size_t li = 0; li < lanes_.size() && !pending_.empty(); ++li) {
            Lane& lane = lanes_[li];
            if (lane.busy) continue;
            // form the batch: oldest ticket's curve, arrival order, within limits
            const Request& head = reqs_[pending_.front()];
            const int curve = head.curve;
            std::vector<int64_t> take;
            std::vector<uint32_t> sizes;
            uint64_t elems = 0;
            if (head.solo) {
                take.push_back(pending_.front());
                sizes.push_back(head.n);
            } else {
                for (int64_t tk : pending_) {
                    const Request& r = reqs_[tk];
                    if (r.curve != curve || r.solo) continue;
                    if (!take.empty() && (take.size() >= cfg_.max_batch || elems + r.n > cfg_.max_batch_elems))
                        continue;
                    take.push_back(tk);
                    sizes.push_back(r.n);
                    elems += r.n;
                }
            }
            std::vector<size_t> order(take.size());
            for (size_t k = 0; k < order.size(); ++k) order[k] = k;
            std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return sizes[a] < sizes[b]; });
            std::vector<int64_t> tickets;
            std::vector<uint32_t> sorted;
            for (size_t k : order) {
                tickets.push_back(take[k]);
                sorted.push_back(sizes[k]);
            }
            const uint64_t fp = exec_.footprint(curve, sorted);
            if (inflight_ > 0 && inflight_ + fp > cfg_.hbm_budget) {
                ++stats_.budget_waits;
                break;
            }
            // admitted: leave the queue, account, then stage and launch unlocked
            for (int64_t tk : take) pending_.erase(std::find(pending_.begin(), pending_.end(), tk));
            lane.busy = true;
            lane.curve = curve;
            lane.bytes = fp;
            lane.tickets = tickets;
            inflight_ += fp;
            stats_.peak_inflight_bytes = std::max<uint64_t>(stats_.peak_inflight_bytes, inflight_);
            stats_.max_batch_requests = std::max<uint64_t>(stats_.max_batch_requests, tickets.size());
            ++stats_.batches;
            std::vector<std::pair<const uint8_t*, uint32_t>> src;
            for (int64_t tk : tickets) src.emplace_back(reqs_[tk].evals, reqs_[tk].n);
            lk.unlock();
            bool ok = true;
            try {
                exec_.launch(li, curve, src, sorted);
            } catch (...) {
                ok = false;
            }
            lk.lock();
            if (!ok) {
                fail(lane, kServeErrBackend);
                ++stats_.launch_failures;
                release(lane);
                cv_done_.notify_all();
            }
            any = true;
        }
        return any;
    }

    void run() {
        std::unique_lock<std::mutex> lk(mu_);
        for (;;) {
            const bool done_some = finish_lanes();
            const bool sent_some = dispatch(lk);
            bool busy = false;
            for (const Lane& lane : lanes_) busy = busy || lane.busy;
            if (stop_ && pending_.empty() && !busy) break;
            if (done_some || sent_some) continue;
            if (busy) cv_sched_.wait_for(lk, std::chrono::microseconds(20));
            else cv_sched_.wait(lk, [&] { return stop_ || !pending_.empty(); });
        }
    }

    Exec& exec_;
    kzg_server_config cfg_;
    size_t max_n_[3] = {0, 0, 0};
    std::mutex mu_;
    std::condition_variable cv_sched_, cv_done_;
    std::deque<int64_t> pending_;
    std::map<int64_t, Request> reqs_;  // node-based: references stay valid across inserts
    int64_t next_ticket_ = 0;
    std::vector<Lane> lanes_;
    uint64_t inflight_ = 0;
    bool stop_ = false;
    kzg_server_stats stats_{};
    std::chrono::steady_clock::time_point t0_ = std::chrono::steady_clock::now();
    std::thread sched_;  // last: started after every member above is ready
};

}  // namespace kz
