// Batched commitments: many polynomials of one curve in one launch sequence
// (one inverse-NTT launch per distinct size, one batched MSM over the SRS
// prefix).  Used by kzg_commit_batch and by the serving loop (serve.hip).
#include <algorithm>
#include <cstdlib>
#include <cstring>

#include "engine.h"
#include "kinet/crypto/kzg.h"
#include "kzg_internal.h"
#include "ntt.h"
#include "scan.h"

namespace kz {

CommitLane::CommitLane(bool own_stream) : owns_stream(own_stream) {
    if (own_stream) KZ_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    KZ_HIP(hipEventCreateWithFlags(&done, hipEventDisableTiming));
}

CommitLane::~CommitLane() {
    if (done) (void)hipEventDestroy(done);
    if (owns_stream && stream) (void)hipStreamDestroy(stream);
}

size_t commit_batch_in_bytes(const std::vector<uint32_t>& n) {
    size_t total = 0;
    for (uint32_t v : n) total += v;
    return 32 * total;
}

size_t commit_batch_out_bytes(int curve, size_t count) { return count * msm_out_bytes(curve) + 16; }

// Batches of one shape are captured into a graph from this use on (the
// capture and instantiation cost about as much as several direct launch
// sequences, so only recurring shapes pay it back).  The environment
// variable KINET_HIP_COMMIT_GRAPH_MIN_USES overrides it; 0 never captures.
static uint32_t graph_min_uses() {
    static const uint32_t v = [] {
        const char* env = std::getenv("KINET_HIP_COMMIT_GRAPH_MIN_USES");
        if (!env || !*env) return 4u;
        char* end = nullptr;
        const unsigned long x = std::strtoul(env, &end, 10);
        return (end && *end == 0 && x < 1000000ul) ? (uint32_t)x : 4u;
    }();
    return v;
}

void commit_batch_launch(CommitLane& lane, const kzg_srs* srs, const std::vector<uint32_t>& n) {
    const int curve = srs->curve;
    const size_t count = n.size();
    size_t total = 0;
    for (uint32_t v : n) total += v;
    hipStream_t s = lane.stream;
    // everything that allocates happens before any launch (a capture may follow)
    uint32_t* ev = static_cast<uint32_t*>(lane.evals.ensure(32 * total));
    uint32_t* co = static_cast<uint32_t*>(lane.coeffs.ensure(32 * total));
    uint32_t* tmp = static_cast<uint32_t*>(lane.tmp.ensure(32 * total));
    const size_t rec = msm_out_bytes(curve);
    uint8_t* out = static_cast<uint8_t*>(lane.out.ensure(commit_batch_out_bytes(curve, count)));
    void* pin_out = lane.pin_out.ensure(commit_batch_out_bytes(curve, count));
    int* err = reinterpret_cast<int*>(out + count * rec);
    const int field = curve == 2 ? 0 : 1;
    for (size_t i = 0; i < count; ++i)
        if (i == 0 || n[i] != n[i - 1]) ntt_prepare(field, log2_exact(n[i]), true, s);
    // A lone job of at least 4096 evaluations runs the single commit's MSM
    // (srs_msm: a window table of the SRS when it has one, else the GLV split
    // when its points are known to lie in G1, else plain windows), which beat
    // the batched form for one job: the same commitment.  Not captured: that
    // pipeline may still grow its workspace on the first run of a size.
    if (count == 1 && n[0] >= 4096) {
        device_fill_u32(reinterpret_cast<uint32_t*>(err), 0, 1, s);
        KZ_HIP(hipMemcpyAsync(ev, lane.pin_in.ptr, 32 * total, hipMemcpyHostToDevice, s));
        ntt_device_batch(field, log2_exact(n[0]), true, ev, co, tmp, 1, err, s);
        srs_msm(srs, co, n[0], 0, lane.ws, s, out);
        KZ_HIP(hipMemcpyAsync(pin_out, out, count * rec + sizeof(int), hipMemcpyDeviceToHost, s));
        ++lane.direct_launches;
        KZ_HIP(hipEventRecord(lane.done, s));
        return;
    }
    // Plans are cached per (curve, SRS size and window tables, job sizes).
    std::vector<uint32_t> kv;
    kv.push_back((uint32_t)srs->n);
    kv.push_back((uint32_t)srs->table_c.size());
    for (int c : srs->table_c) kv.push_back((uint32_t)c);
    kv.insert(kv.end(), n.begin(), n.end());
    auto key = std::make_pair(curve, kv);
    auto it = lane.plans.find(key);
    if (it == lane.plans.end()) {
        if (lane.plans.size() >= 1024) lane.plans.clear();
        std::vector<MsmJob> specs(count);
        uint64_t sc = 0;
        for (size_t i = 0; i < count; ++i) {
            specs[i] = MsmJob{};
            specs[i].n = n[i];
            specs[i].sc_base = sc;
            sc += n[i];
            if (!srs->table_c.empty()) {
                // every job indexes the tables buffer: a table within 2 bits
                // of the width a single commit of this size takes a table at
                // (msm_table_window_bits), else one within 2 bits of the
                // built-in width, else ordinary windows at the job's batched
                // width over row 0 (the points) of the first table.  (The
                // tuned batched widths are table-free optima: matched against
                // them, BN254 jobs of 2^19-2^20 points took a c = 12 table
                // where c = 15 was there, and the 70/30 serving mix ran at
                // 1920-1950 commits/s against 2590-2600 with this rule.)
                const int want = msm_window_bits(curve, n[i], kMsmBatched);
                int k = srs_table_near(srs, msm_table_window_bits(n[i]), 2);
                if (k < 0) k = srs_table_near(srs, msm_default_window_bits(n[i]), 2);
                if (k >= 0) {
                    specs[i].c = (uint32_t)srs->table_c[k];
                    specs[i].table = (uint32_t)srs->n;
                    specs[i].tab_base = srs->table_base[k];
                } else {
                    specs[i].c = (uint32_t)want;
                    specs[i].tab_base = srs->table_base[0];
                }
            }
        }
        auto lp = std::make_unique<LanePlan>();
        msm_batch_plan(curve, specs.data(), count, 0, lp->plan);
        it = lane.plans.emplace(key, std::move(lp)).first;
    }
    LanePlan& lp = *it->second;
    msm_batch_reserve(lp.plan, lane.ws);
    const uint32_t* pts = srs->table_c.empty() ? srs->pts.as<uint32_t>() : srs->tables.as<uint32_t>();
    const uint8_t* inf = srs->inf.as<uint8_t>();
    auto launch = [&] {
        device_fill_u32(reinterpret_cast<uint32_t*>(err), 0, 1, s);
        KZ_HIP(hipMemcpyAsync(ev, lane.pin_in.ptr, 32 * total, hipMemcpyHostToDevice, s));
        size_t off = 0;
        for (size_t i = 0; i < count;) {
            size_t j = i;
            while (j < count && n[j] == n[i]) ++j;
            ntt_device_batch(field, log2_exact(n[i]), true, ev + off * 8, co + off * 8, tmp + off * 8, j - i, err, s);
            off += (size_t)n[i] * (j - i);
            i = j;
        }
        msm_batch_run(lp.plan, pts, inf, co, lane.ws, s, out);
        KZ_HIP(hipMemcpyAsync(pin_out, out, count * rec + sizeof(int), hipMemcpyDeviceToHost, s));
    };
    ++lp.uses;
    const void* bufs[LanePlan::kAddrs] = {ev, co, tmp, out, lane.pin_in.ptr, pin_out, pts, inf};
    lane.ws.addresses(bufs + 8);
    const uint32_t min_uses = graph_min_uses();
    bool use_graph = min_uses > 0 && lp.uses >= min_uses && !lp.no_graph;
    if (use_graph && !(lp.exec && std::equal(bufs, bufs + LanePlan::kAddrs, lp.bufs))) {
        lp.drop_graph();
        KZ_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        try {
            launch();
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(s, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        KZ_HIP(hipStreamEndCapture(s, &lp.graph));
        if (hipGraphInstantiate(&lp.exec, lp.graph, nullptr, nullptr, 0) != hipSuccess) {
            lp.drop_graph();
            lp.no_graph = true;
            use_graph = false;
        } else {
            std::copy(bufs, bufs + LanePlan::kAddrs, lp.bufs);
        }
    }
    if (use_graph) {
        KZ_HIP(hipGraphLaunch(lp.exec, s));
        ++lane.graph_launches;
    } else {
        launch();
        ++lane.direct_launches;
    }
    KZ_HIP(hipEventRecord(lane.done, s));
}

static CommitLane& engine_lane() {
    Engine& e = engine();
    static CommitLane* lane = [&e] {
        CommitLane* l = new CommitLane(false);
        l->stream = e.stream;
        return l;
    }();
    return *lane;
}

}  // namespace kz

using namespace kz;

extern "C" {

// Jobs that share an SRS handle run as one batch (sizes sorted so that equal
// sizes share inverse-NTT launches); different handles run one after another.
int kzg_commit_batch(const kzg_srs* const* srs, const uint8_t* const* evals, const size_t* n, size_t count,
                     uint8_t* commits_out) {
    if (count == 0) return KZ_CRYPTO_OK;
    if (!srs || !evals || !n || !commits_out || !srs[0]) return KZ_CRYPTO_ERR_INPUT;
    const int curve = srs[0]->curve;
    const size_t cb = compressed_bytes(curve);
    for (size_t i = 0; i < count; ++i) {
        if (!srs[i] || srs[i]->curve != curve || !evals[i]) return KZ_CRYPTO_ERR_INPUT;
        const int lg = log2_exact(n[i]);
        if (n[i] == 0 || lg < 0 || lg > (curve == 2 ? 32 : 28) || n[i] > srs[i]->n || n[i] > (1u << 26))
            return KZ_CRYPTO_ERR_INPUT;
    }
    try {
        Engine& e = engine();
        std::lock_guard<std::mutex> g(e.mu);
        CommitLane& lane = engine_lane();
        std::map<const kzg_srs*, std::vector<size_t>> groups;
        for (size_t i = 0; i < count; ++i) groups[srs[i]].push_back(i);
        std::vector<uint8_t> result(cb * count);
        for (auto& grp : groups) {
            std::vector<size_t>& idx = grp.second;
            std::stable_sort(idx.begin(), idx.end(), [n](size_t a, size_t b) { return n[a] < n[b]; });
            std::vector<uint32_t> sizes;
            for (size_t i : idx) sizes.push_back((uint32_t)n[i]);
            uint8_t* pin = static_cast<uint8_t*>(lane.pin_in.ensure(commit_batch_in_bytes(sizes)));
            lane.pin_out.ensure(commit_batch_out_bytes(curve, idx.size()));
            size_t off = 0;
            for (size_t i : idx) {
                std::memcpy(pin + off, evals[i], 32 * n[i]);
                off += 32 * n[i];
            }
            commit_batch_launch(lane, grp.first, sizes);
            KZ_HIP(hipEventSynchronize(lane.done));
            const uint8_t* po = lane.pin_out.as<uint8_t>();
            const size_t rec = msm_out_bytes(curve);
            int err;
            std::memcpy(&err, po + idx.size() * rec, sizeof(int));
            if (err) return KZ_CRYPTO_ERR_INPUT;
            for (size_t k = 0; k < idx.size(); ++k)
                std::memcpy(result.data() + cb * idx[k], po + k * rec + point_wire_bytes(curve), cb);
        }
        std::memcpy(commits_out, result.data(), cb * count);
        return KZ_CRYPTO_OK;
    } catch (const std::exception& ex) {
        set_last_error(ex.what());
        return KZ_CRYPTO_ERR_BACKEND;
    }
}

}  // extern "C"
