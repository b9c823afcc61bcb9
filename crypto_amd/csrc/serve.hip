// KZG commitment serving loop (include/kinet/crypto/kzg_serve.h): the
// scheduler of serve_core.h over the HIP executor below, whose lanes are
// CommitLanes (a stream, its buffers and page-locked staging) running the
// batched commit pipeline of kzg_batch.hip.
#include <memory>

#include "engine.h"
#include "kinet/crypto/kzg_serve.h"
#include "kzg_internal.h"
#include "serve_core.h"

namespace kz {

namespace {

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

// Executor of ServeCore on the device: one CommitLane per lane.
struct HipExec {
    const kzg_srs* srs[3] = {nullptr, nullptr, nullptr};
    std::vector<std::unique_ptr<CommitLane>> cl;
    std::vector<int> curve;

    size_t lanes() const { return cl.size(); }
    uint64_t footprint(int c, const std::vector<uint32_t>& sizes) { return batch_footprint(c, sizes); }
    void launch(size_t li, int c, const std::vector<std::pair<const uint8_t*, uint32_t>>& src,
                const std::vector<uint32_t>& sizes) {
        try {
            CommitLane& lane = *cl[li];
            uint8_t* pin = static_cast<uint8_t*>(lane.pin_in.ensure(commit_batch_in_bytes(sizes)));
            lane.pin_out.ensure(commit_batch_out_bytes(c, sizes.size()));
            size_t off = 0;
            for (auto& p : src) {
                std::memcpy(pin + off, p.first, 32 * (size_t)p.second);
                off += 32 * (size_t)p.second;
            }
            curve[li] = c;
            commit_batch_launch(lane, srs[c], sizes);
        } catch (const std::exception& ex) {
            set_last_error(ex.what());
            throw;
        }
    }
    int poll(size_t li) {
        const hipError_t q = hipEventQuery(cl[li]->done);
        if (q == hipErrorNotReady) return 0;
        if (q == hipSuccess) return 1;
        set_last_error(hipGetErrorString(q));
        return -1;
    }
    bool error_flag(size_t li, size_t count) {
        int err = 0;
        std::memcpy(&err, cl[li]->pin_out.as<uint8_t>() + count * msm_out_bytes(curve[li]), sizeof(int));
        return err != 0;
    }
    const uint8_t* result(size_t li, size_t k) {
        return cl[li]->pin_out.as<uint8_t>() + k * msm_out_bytes(curve[li]) + point_wire_bytes(curve[li]);
    }
};

}  // namespace

}  // namespace kz

using namespace kz;

struct kzg_server {
    HipExec exec;                                   // destroyed after the scheduler below
    std::unique_ptr<ServeCore<HipExec>> core;
};

extern "C" {

int kzg_server_create(const kzg_srs* srs_bn254, const kzg_srs* srs_bls12_381, const kzg_server_config* cfg,
                      kzg_server** out) {
    if (!out || (!srs_bn254 && !srs_bls12_381)) return KZ_CRYPTO_ERR_INPUT;
    if ((srs_bn254 && srs_bn254->curve != 1) || (srs_bls12_381 && srs_bls12_381->curve != 2))
        return KZ_CRYPTO_ERR_INPUT;
    *out = nullptr;
    try {
        (void)engine();
        kzg_server_config c{};
        if (cfg) c = *cfg;
        if (!c.lanes) c.lanes = 3;
        if (!c.max_batch) c.max_batch = 128;
        if (!c.max_batch_elems) {
            // the largest admissible request (a power of two within an SRS),
            // at least 2^20: on the 70/30 mix of BLS12-381 4096-point and
            // BN254 2^14..2^20 commits, batches of up to 2^20 evaluations
            // served 2310-2380 commits/s closed loop against 1860-1990 for
            // 2^21, with p99 11.9-13.3 ms against 14.5-15.1 ms at 1560/s
            size_t top = 0;
            for (const kzg_srs* h : {srs_bn254, srs_bls12_381})
                if (h) top = std::max(top, h->n);
            uint64_t elems = 1ull << 20;
            while (elems * 2 <= top) elems *= 2;
            c.max_batch_elems = elems;
        }
        if (!c.hbm_budget) c.hbm_budget = 8ull << 30;
        if (c.lanes > 16 || c.max_batch > 65535) return KZ_CRYPTO_ERR_INPUT;
        auto s = std::make_unique<kzg_server>();
        s->exec.srs[1] = srs_bn254;
        s->exec.srs[2] = srs_bls12_381;
        for (uint32_t k = 0; k < c.lanes; ++k) s->exec.cl.push_back(std::make_unique<CommitLane>(true));
        s->exec.curve.assign(c.lanes, 0);
        const size_t max_n[3] = {0, srs_bn254 ? srs_bn254->n : 0, srs_bls12_381 ? srs_bls12_381->n : 0};
        s->core = std::make_unique<ServeCore<HipExec>>(s->exec, c, max_n);
        *out = s.release();
        return KZ_CRYPTO_OK;
    } catch (const std::exception& ex) {
        set_last_error(ex.what());
        return KZ_CRYPTO_ERR_BACKEND;
    }
}

void kzg_server_destroy(kzg_server* s) { delete s; }  // the scheduler finishes admitted work first

int64_t kzg_server_submit(kzg_server* s, uint32_t curve, const uint8_t* evals, size_t n) {
    if (!s) return KZ_CRYPTO_ERR_INPUT;
    return s->core->submit(curve, evals, n);
}

int kzg_server_wait(kzg_server* s, int64_t ticket, uint8_t* commit_out, uint64_t* t_submit_ns,
                    uint64_t* t_done_ns) {
    if (!s || !commit_out) return KZ_CRYPTO_ERR_INPUT;
    return s->core->wait(ticket, commit_out, t_submit_ns, t_done_ns);
}

int kzg_server_get_stats(const kzg_server* s, kzg_server_stats* out) {
    if (!s || !out) return KZ_CRYPTO_ERR_INPUT;
    *out = const_cast<kzg_server*>(s)->core->stats();
    return KZ_CRYPTO_OK;
}

}  // extern "C"
