// KZG commitment serving loop (include/kinet/crypto/kzg_serve.h).
//
// Submitters append tickets to a FIFO.  One scheduler thread owns the lanes
// (CommitLane: a stream plus its buffers) and loops:
//   1. completions: a busy lane whose done event has fired hands its results
//      to the tickets and returns its footprint to the HBM budget;
//   2. dispatch: while a lane is free, it takes the oldest pending ticket and,
//      in arrival order, every further pending ticket of the same curve that
//      fits the batch limits (continuous batching: whatever is waiting when a
//      lane frees up rides together); the batch is admitted only if its
//      device footprint fits the budget left by the batches in flight, else
//      it waits for a completion (admission control).  The lane stages the
//      evaluations into its page-locked buffer and launches the batch.
// A batch whose error flag is set (some evaluation >= r) is re-run one
// request per batch, so only the offending request fails.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <thread>
#include <map>

#include "engine.h"
#include "kinet/crypto/kzg_serve.h"
#include "kzg_internal.h"

namespace kz {

namespace {

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
    std::unique_ptr<CommitLane> cl;
    bool busy = false;
    int curve = 0;
    uint64_t bytes = 0;
    std::vector<int64_t> tickets;  // in batch (size-sorted) order
};

// Device bytes a batch touches: evaluations, coefficients and NTT scratch,
// plus the MSM's sorted entries, carries, bucket sums and reduction buffers.
uint64_t batch_footprint(int curve, const std::vector<uint32_t>& n) {
    const uint64_t rec = (uint64_t)xyzz_words(curve) * 4;
    uint64_t bytes = 0;
    for (uint32_t v : n) {
        const int c = msm_window_bits(curve, v, kMsmBatched);
        const uint64_t W = (uint64_t)msm_num_windows(c), B = 1ull << (c - 1);
        uint64_t parts = 1;  // the batch sort's parts per window
        while (parts < 64 && parts * 2 * 8192 <= v) parts *= 2;
        bytes += 3ull * 32 * v;                     // evals, coeffs, scratch
        bytes += (uint64_t)v * W * (4 + rec / 32);  // entries + carries
        bytes += (uint64_t)v * W * 2;               // the sort's digit codes
        bytes += W * parts * B * 8;                 // the sort's per-part counts and prefixes
        bytes += W * B * (12 + rec + rec / 8);      // counts/offsets/cursor, bucket sums, segments
    }
    return bytes;
}

}  // namespace

}  // namespace kz

using namespace kz;

struct kzg_server {
    const kzg_srs* srs[3] = {nullptr, nullptr, nullptr};
    kzg_server_config cfg{};
    std::mutex mu;
    std::condition_variable cv_sched, cv_done;
    std::deque<int64_t> pending;
    std::map<int64_t, Request> reqs;  // node-based: references stay valid across inserts
    int64_t next_ticket = 0;
    std::vector<Lane> lanes;
    uint64_t inflight = 0;
    bool stop = false;
    kzg_server_stats stats{};
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    std::thread sched;

    uint64_t now_ns() const {
        return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now() - t0)
            .count();
    }
    void run();
    bool finish_lanes();                          // under mu
    bool dispatch(std::unique_lock<std::mutex>& lk);  // under mu (released while staging)
};

bool kzg_server::finish_lanes() {
    bool any = false;
    for (Lane& lane : lanes) {
        if (!lane.busy) continue;
        const hipError_t q = hipEventQuery(lane.cl->done);
        if (q == hipErrorNotReady) continue;
        any = true;
        const uint64_t t = now_ns();
        const size_t rec = msm_out_bytes(lane.curve), cb = compressed_bytes(lane.curve);
        const uint8_t* po = lane.cl->pin_out.as<uint8_t>();
        int err = 0;
        if (q == hipSuccess) std::memcpy(&err, po + lane.tickets.size() * rec, sizeof(int));
        if (q != hipSuccess || (err && lane.tickets.size() == 1)) {
            for (int64_t tk : lane.tickets) {
                Request& r = reqs[tk];
                r.status = q != hipSuccess ? KZ_CRYPTO_ERR_BACKEND : KZ_CRYPTO_ERR_INPUT;
                r.t_done = t;
                ++stats.failed;
            }
            if (q != hipSuccess) set_last_error(hipGetErrorString(q));
        } else if (err) {
            // some evaluation >= r: re-run each request alone, in order, at the front
            for (auto it = lane.tickets.rbegin(); it != lane.tickets.rend(); ++it) {
                reqs[*it].solo = true;
                pending.push_front(*it);
            }
        } else {
            for (size_t k = 0; k < lane.tickets.size(); ++k) {
                Request& r = reqs[lane.tickets[k]];
                std::memcpy(r.commit, po + k * rec + point_wire_bytes(lane.curve), cb);
                r.status = 1;
                r.t_done = t;
                ++stats.completed;
            }
        }
        inflight -= lane.bytes;
        lane.busy = false;
        lane.tickets.clear();
    }
    if (any) cv_done.notify_all();
    return any;
}

bool kzg_server::dispatch(std::unique_lock<std::mutex>& lk) {
    bool any = false;
    for (size_t li = 0; li < lanes.size() && !pending.empty(); ++li) {
        Lane& lane = lanes[li];
        if (lane.busy) continue;
        // form the batch: oldest ticket's curve, arrival order, within limits
        const Request& head = reqs[pending.front()];
        const int curve = head.curve;
        std::vector<int64_t> take;
        std::vector<uint32_t> sizes;
        uint64_t elems = 0;
        if (head.solo) {
            take.push_back(pending.front());
            sizes.push_back(head.n);
        } else {
            for (int64_t tk : pending) {
                const Request& r = reqs[tk];
                if (r.curve != curve || r.solo) continue;
                if (!take.empty() && (take.size() >= cfg.max_batch || elems + r.n > cfg.max_batch_elems)) continue;
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
        const uint64_t fp = batch_footprint(curve, sorted);
        if (inflight > 0 && inflight + fp > cfg.hbm_budget) {
            ++stats.budget_waits;
            break;
        }
        // admitted: leave the queue, account, then stage and launch unlocked
        for (int64_t tk : take) pending.erase(std::find(pending.begin(), pending.end(), tk));
        lane.busy = true;
        lane.curve = curve;
        lane.bytes = fp;
        lane.tickets = tickets;
        inflight += fp;
        stats.peak_inflight_bytes = std::max<uint64_t>(stats.peak_inflight_bytes, inflight);
        stats.max_batch_requests = std::max<uint64_t>(stats.max_batch_requests, tickets.size());
        ++stats.batches;
        std::vector<std::pair<const uint8_t*, uint32_t>> src;
        for (int64_t tk : tickets) src.emplace_back(reqs[tk].evals, reqs[tk].n);
        const kzg_srs* srs = this->srs[curve];
        lk.unlock();
        bool ok = true;
        try {
            uint8_t* pin = static_cast<uint8_t*>(lane.cl->pin_in.ensure(commit_batch_in_bytes(sorted)));
            lane.cl->pin_out.ensure(commit_batch_out_bytes(curve, sorted.size()));
            size_t off = 0;
            for (auto& p : src) {
                std::memcpy(pin + off, p.first, 32 * (size_t)p.second);
                off += 32 * (size_t)p.second;
            }
            commit_batch_launch(*lane.cl, srs, sorted);
        } catch (const std::exception& ex) {
            set_last_error(ex.what());
            ok = false;
        }
        lk.lock();
        if (!ok) {
            const uint64_t t = now_ns();
            for (int64_t tk : lane.tickets) {
                reqs[tk].status = KZ_CRYPTO_ERR_BACKEND;
                reqs[tk].t_done = t;
                ++stats.failed;
            }
            inflight -= lane.bytes;
            lane.busy = false;
            lane.tickets.clear();
            cv_done.notify_all();
        }
        any = true;
    }
    return any;
}

void kzg_server::run() {
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        const bool done_some = finish_lanes();
        const bool sent_some = dispatch(lk);
        bool busy = false;
        for (const Lane& lane : lanes) busy = busy || lane.busy;
        if (stop && pending.empty() && !busy) break;
        if (done_some || sent_some) continue;
        if (busy) cv_sched.wait_for(lk, std::chrono::microseconds(20));
        else cv_sched.wait(lk, [&] { return stop || !pending.empty(); });
    }
}

extern "C" {

int kzg_server_create(const kzg_srs* srs_bn254, const kzg_srs* srs_bls12_381, const kzg_server_config* cfg,
                      kzg_server** out) {
    if (!out || (!srs_bn254 && !srs_bls12_381)) return KZ_CRYPTO_ERR_INPUT;
    if ((srs_bn254 && srs_bn254->curve != 1) || (srs_bls12_381 && srs_bls12_381->curve != 2))
        return KZ_CRYPTO_ERR_INPUT;
    *out = nullptr;
    try {
        (void)engine();
        auto s = std::make_unique<kzg_server>();
        s->srs[1] = srs_bn254;
        s->srs[2] = srs_bls12_381;
        if (cfg) s->cfg = *cfg;
        if (!s->cfg.lanes) s->cfg.lanes = 3;
        if (!s->cfg.max_batch) s->cfg.max_batch = 128;
        if (!s->cfg.max_batch_elems) s->cfg.max_batch_elems = 1ull << 21;
        if (!s->cfg.hbm_budget) s->cfg.hbm_budget = 8ull << 30;
        if (s->cfg.lanes > 16 || s->cfg.max_batch > 65535) return KZ_CRYPTO_ERR_INPUT;
        s->lanes.resize(s->cfg.lanes);
        for (Lane& lane : s->lanes) lane.cl = std::make_unique<CommitLane>(true);
        kzg_server* raw = s.get();
        s->sched = std::thread([raw] { raw->run(); });
        *out = s.release();
        return KZ_CRYPTO_OK;
    } catch (const std::exception& ex) {
        set_last_error(ex.what());
        return KZ_CRYPTO_ERR_BACKEND;
    }
}

void kzg_server_destroy(kzg_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
    }
    s->cv_sched.notify_all();
    if (s->sched.joinable()) s->sched.join();
    delete s;
}

int64_t kzg_server_submit(kzg_server* s, uint32_t curve, const uint8_t* evals, size_t n) {
    if (!s || !evals || (curve != 1 && curve != 2) || !s->srs[curve]) return KZ_CRYPTO_ERR_INPUT;
    const int lg = log2_exact(n);
    if (n == 0 || lg < 0 || n > s->srs[curve]->n || n > s->cfg.max_batch_elems || n > (1u << 26))
        return KZ_CRYPTO_ERR_INPUT;
    int64_t tk;
    {
        std::lock_guard<std::mutex> g(s->mu);
        if (s->stop) return KZ_CRYPTO_ERR_INPUT;
        tk = s->next_ticket++;
        Request& r = s->reqs[tk];
        r.curve = (int)curve;
        r.evals = evals;
        r.n = (uint32_t)n;
        r.t_submit = s->now_ns();
        s->pending.push_back(tk);
        ++s->stats.submitted;
    }
    s->cv_sched.notify_one();
    return tk;
}

int kzg_server_wait(kzg_server* s, int64_t ticket, uint8_t* commit_out, uint64_t* t_submit_ns,
                    uint64_t* t_done_ns) {
    if (!s || !commit_out) return KZ_CRYPTO_ERR_INPUT;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return KZ_CRYPTO_ERR_INPUT;
    s->cv_done.wait(lk, [&] { return it->second.status != 0; });  // std::map: `it` survives inserts
    const Request r = it->second;
    s->reqs.erase(it);
    lk.unlock();
    if (t_submit_ns) *t_submit_ns = r.t_submit;
    if (t_done_ns) *t_done_ns = r.t_done;
    if (r.status < 0) return r.status;
    std::memcpy(commit_out, r.commit, compressed_bytes(r.curve));
    return KZ_CRYPTO_OK;
}

int kzg_server_get_stats(const kzg_server* s, kzg_server_stats* out) {
    if (!s || !out) return KZ_CRYPTO_ERR_INPUT;
    std::lock_guard<std::mutex> g(const_cast<kzg_server*>(s)->mu);
    *out = s->stats;
    return KZ_CRYPTO_OK;
}

}  // extern "C"
