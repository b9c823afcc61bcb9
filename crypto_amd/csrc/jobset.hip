// Job sets: a fixed set of MSMs and Fr NTTs of uneven sizes run as one HIP
// graph replay.  The MSMs form one batch plan (every pipeline stage launched
// once for all jobs); the NTTs run one batched launch sequence per size
// group.  The launch sequence is captured once per set of argument pointers
// and stream, with the NTT groups on a second stream forked inside the
// capture, so the graph has two branches: NTT passes fill the device while
// the MSM's bucket reduction and window combine (latency-bound, few waves)
// run, and one hipGraphLaunch starts the whole set.
//
// No counterpart in the reference: its GPU paths launch one blob at a time
// (kzg/gpu/cuda/kzg_driver_cuda.cpp:44-58).
#include <memory>
#include <vector>

#include "engine.h"
#include "kinet/crypto/hip_runtime.h"
#include "kinet/gpukit/gpukit.h"
#include "ntt.h"
#include "scan.h"

namespace kz {
namespace {

struct SetKey {
    const void* p[8] = {};
    hipStream_t stream = nullptr;
    bool operator==(const SetKey& o) const {
        for (int i = 0; i < 8; ++i)
            if (p[i] != o.p[i]) return false;
        return stream == o.stream;
    }
};

}  // namespace
}  // namespace kz

struct kinet_job_set {
    int curve = 0, field = 0;
    bool inverse = false;
    bool has_msm = false;
    kz::MsmBatchPlan plan;
    kz::MsmWorkspace ws;
    std::vector<kinet_ntt_group> groups;
    hipStream_t side = nullptr;  // the NTT branch inside the capture
    hipEvent_t fork = nullptr, join = nullptr;
    kz::SetKey key;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;

    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
    ~kinet_job_set() {
        drop_graph();
        if (fork) (void)hipEventDestroy(fork);
        if (join) (void)hipEventDestroy(join);
        if (side) (void)hipStreamDestroy(side);
    }
};

using namespace kz;

extern "C" {

int kinet_hip_job_set_create(uint32_t curve, const size_t* msm_n, const uint64_t* sc_off, const uint32_t* pt_off,
                             size_t msm_count, uint32_t field, int ntt_inverse, const kinet_ntt_group* groups,
                             size_t ngroups, kinet_job_set** out) {
    if (!out || (msm_count > 0 && !msm_n) || (ngroups > 0 && !groups)) return GPUKIT_ERR_NULL_ARG;
    *out = nullptr;
    if (msm_count > 0 && curve != 1 && curve != 2) return GPUKIT_ERR_BACKEND;
    if (ngroups > 0 && field > 1) return GPUKIT_ERR_BACKEND;
    if (msm_count > (1u << 24) || msm_count + ngroups == 0) return GPUKIT_ERR_BAD_SIZE;
    for (size_t g = 0; g < ngroups; ++g)
        if (groups[g].count == 0 || groups[g].count > 65535 || groups[g].log_n > (uint32_t)kNttMaxLog ||
            groups[g].log_n > (field == 0 ? 32u : 28u))
            return GPUKIT_ERR_BAD_SIZE;
    try {
        Engine& e = engine();
        auto set = std::make_unique<kinet_job_set>();
        set->curve = (int)curve;
        set->field = (int)field;
        set->inverse = ntt_inverse != 0;
        set->groups.assign(groups, groups + ngroups);
        if (msm_count > 0) {
            std::vector<MsmJob> specs(msm_count);
            uint64_t next_sc = 0;
            for (size_t j = 0; j < msm_count; ++j) {
                if (msm_n[j] == 0 || msm_n[j] >= ((size_t)1 << 26)) return GPUKIT_ERR_BAD_SIZE;
                specs[j] = MsmJob{};
                specs[j].n = (uint32_t)msm_n[j];
                specs[j].sc_base = sc_off ? sc_off[j] : next_sc;
                specs[j].pt_base = pt_off ? pt_off[j] : 0;
                next_sc += msm_n[j];
            }
            msm_batch_plan((int)curve, specs.data(), msm_count, 0, set->plan);
            msm_batch_reserve(set->plan, set->ws);
            set->has_msm = true;
        }
        // twiddle tables now: the captured sequence may only launch kernels
        for (auto& g : set->groups) ntt_prepare((int)field, (int)g.log_n, set->inverse, e.stream);
        KZ_HIP(hipStreamSynchronize(e.stream));
        KZ_HIP(hipStreamCreateWithFlags(&set->side, hipStreamNonBlocking));
        KZ_HIP(hipEventCreateWithFlags(&set->fork, hipEventDisableTiming));
        KZ_HIP(hipEventCreateWithFlags(&set->join, hipEventDisableTiming));
        *out = set.release();
        return GPUKIT_OK;
    } catch (const std::invalid_argument& ex) {
        set_last_error(ex.what());
        return GPUKIT_ERR_BAD_SIZE;
    } catch (const std::exception& ex) {
        set_last_error(ex.what());
        return GPUKIT_ERR_BACKEND;
    }
}

void kinet_hip_job_set_destroy(kinet_job_set* set) { delete set; }

int kinet_hip_job_set_run(kinet_job_set* set, const void* d_scalars, const void* d_pts, const void* d_inf,
                          void* d_msm_out, const void* d_ntt_in, void* d_ntt_out, void* d_ntt_tmp, int* d_err,
                          void* stream) {
    if (!set || !stream) return GPUKIT_ERR_NULL_ARG;
    if (set->has_msm && (!d_scalars || !d_pts || !d_msm_out)) return GPUKIT_ERR_NULL_ARG;
    if (!set->groups.empty() && (!d_ntt_in || !d_ntt_out || !d_ntt_tmp || !d_err)) return GPUKIT_ERR_NULL_ARG;
    hipStream_t s = static_cast<hipStream_t>(stream);
    try {
        SetKey key;
        const void* ptrs[8] = {d_scalars, d_pts, d_inf, d_msm_out, d_ntt_in, d_ntt_out, d_ntt_tmp, d_err};
        for (int i = 0; i < 8; ++i) key.p[i] = ptrs[i];
        key.stream = s;
        if (!set->exec || !(set->key == key)) {
            set->drop_graph();
            KZ_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            try {
                const bool ntt = !set->groups.empty();
                if (ntt) {
                    device_fill_u32(reinterpret_cast<uint32_t*>(d_err), 0, 1, s);
                    KZ_HIP(hipEventRecord(set->fork, s));
                    KZ_HIP(hipStreamWaitEvent(set->side, set->fork, 0));
                    for (auto& g : set->groups) {
                        const size_t off = (size_t)g.offset * 8;  // 8 words per element
                        ntt_device_batch(set->field, (int)g.log_n, set->inverse,
                                         static_cast<const uint32_t*>(d_ntt_in) + off,
                                         static_cast<uint32_t*>(d_ntt_out) + off,
                                         static_cast<uint32_t*>(d_ntt_tmp) + off, g.count, d_err, set->side);
                    }
                    KZ_HIP(hipEventRecord(set->join, set->side));
                }
                if (set->has_msm)
                    msm_batch_run(set->plan, static_cast<const uint32_t*>(d_pts), static_cast<const uint8_t*>(d_inf),
                                  static_cast<const uint32_t*>(d_scalars), set->ws, s,
                                  static_cast<uint8_t*>(d_msm_out));
                if (ntt) KZ_HIP(hipStreamWaitEvent(s, set->join, 0));
            } catch (...) {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(s, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            KZ_HIP(hipStreamEndCapture(s, &set->graph));
            KZ_HIP(hipGraphInstantiate(&set->exec, set->graph, nullptr, nullptr, 0));
            set->key = key;
        }
        KZ_HIP(hipGraphLaunch(set->exec, s));
        return GPUKIT_OK;
    } catch (const std::invalid_argument& ex) {
        set_last_error(ex.what());
        return GPUKIT_ERR_BAD_SIZE;
    } catch (const std::exception& ex) {
        set_last_error(ex.what());
        return GPUKIT_ERR_BACKEND;
    }
}

int kinet_hip_job_set_window_bits(const kinet_job_set* set, size_t job) {
    if (!set || !set->has_msm || job >= set->plan.jobs.size()) return -1;
    return (int)set->plan.jobs[job].c;
}

}  // extern "C"
