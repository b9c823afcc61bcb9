// C ABI of batched MSMs: a plan per batch shape (descriptor tables resident on
// the device, own scratch), run by direct launches or by replaying a HIP
// graph captured from the same launch sequence.
#include <memory>

#include "engine.h"
#include "kinet/crypto/hip_runtime.h"
#include "kinet/gpukit/gpukit.h"

namespace kz {

struct GraphKey {
    const void* sc = nullptr;
    const void* pts = nullptr;
    const void* inf = nullptr;
    void* out = nullptr;
    hipStream_t stream = nullptr;
    bool operator==(const GraphKey& o) const {
        return sc == o.sc && pts == o.pts && inf == o.inf && out == o.out && stream == o.stream;
    }
};

}  // namespace kz

struct kinet_msm_batch {
    kz::MsmBatchPlan plan;
    kz::MsmWorkspace ws;
    kz::GraphKey key;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;

    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
    ~kinet_msm_batch() { drop_graph(); }
};

using namespace kz;

extern "C" {

int kinet_hip_msm_batch_create(uint32_t curve, const size_t* n, const uint64_t* sc_off, const uint32_t* pt_off,
                               size_t count, int window_bits, kinet_msm_batch** out) {
    if (!out || !n) return GPUKIT_ERR_NULL_ARG;
    *out = nullptr;
    if (curve != 1 && curve != 2) return GPUKIT_ERR_BACKEND;
    if (count == 0 || count > (1u << 24)) return GPUKIT_ERR_BAD_SIZE;
    if (window_bits != 0 && (window_bits < 5 || window_bits > 18)) return GPUKIT_ERR_BAD_SIZE;
    try {
        (void)engine();
        std::vector<MsmJob> specs(count);
        uint64_t next_sc = 0;
        for (size_t j = 0; j < count; ++j) {
            if (n[j] >= ((size_t)1 << 26)) return GPUKIT_ERR_BAD_SIZE;
            MsmJob& s = specs[j];
            s = MsmJob{};
            s.n = (uint32_t)n[j];
            s.sc_base = sc_off ? sc_off[j] : next_sc;
            s.pt_base = pt_off ? pt_off[j] : 0;
            s.c = (uint32_t)window_bits;
            next_sc += n[j];
        }
        auto b = std::make_unique<kinet_msm_batch>();
        msm_batch_plan((int)curve, specs.data(), count, 0, b->plan);
        msm_batch_reserve(b->plan, b->ws);
        *out = b.release();
        return GPUKIT_OK;
    } catch (const std::invalid_argument& ex) {
        set_last_error(ex.what());
        return GPUKIT_ERR_BAD_SIZE;
    } catch (const std::exception& ex) {
        set_last_error(ex.what());
        return GPUKIT_ERR_BACKEND;
    }
}

void kinet_hip_msm_batch_destroy(kinet_msm_batch* batch) { delete batch; }

int kinet_hip_msm_batch_window_bits(const kinet_msm_batch* batch, size_t job) {
    if (!batch || job >= batch->plan.jobs.size()) return -1;
    return (int)batch->plan.jobs[job].c;
}

int kinet_hip_msm_batch_run(kinet_msm_batch* b, const void* d_scalars, const void* d_pts, const void* d_inf,
                            void* d_out, int use_graph, void* stream) {
    if (!b || !d_scalars || !d_pts || !d_out) return GPUKIT_ERR_NULL_ARG;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (use_graph && !s) return GPUKIT_ERR_NULL_ARG;
    try {
        auto launch = [&] {
            msm_batch_run(b->plan, static_cast<const uint32_t*>(d_pts), static_cast<const uint8_t*>(d_inf),
                          static_cast<const uint32_t*>(d_scalars), b->ws, s, static_cast<uint8_t*>(d_out));
        };
        if (!use_graph) {
            launch();
            return GPUKIT_OK;
        }
        const GraphKey key{d_scalars, d_pts, d_inf, d_out, s};
        if (!b->exec || !(b->key == key)) {
            b->drop_graph();
            KZ_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            try {
                launch();
            } catch (...) {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(s, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            KZ_HIP(hipStreamEndCapture(s, &b->graph));
            KZ_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
            b->key = key;
        }
        KZ_HIP(hipGraphLaunch(b->exec, s));
        return GPUKIT_OK;
    } catch (const std::exception& ex) {
        set_last_error(ex.what());
        return GPUKIT_ERR_BACKEND;
    }
}

int kinet_hip_msm_set_window_table(uint32_t curve, int mode, const int* c_by_log, int count) {
    if (count > 0 && !c_by_log) return GPUKIT_ERR_NULL_ARG;
    try {
        msm_set_window_table((int)curve, mode, c_by_log, count);
        return GPUKIT_OK;
    } catch (const std::exception& ex) {
        set_last_error(ex.what());
        return GPUKIT_ERR_BAD_SIZE;
    }
}

int kinet_hip_msm_window_bits(uint32_t curve, size_t n, int mode) {
    if (!msm_tunable((int)curve, mode)) return -1;
    return msm_window_bits((int)curve, n, mode);
}

}  // extern "C"
