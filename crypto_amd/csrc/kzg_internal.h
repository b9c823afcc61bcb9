// Internal state behind include/kinet/crypto/kzg.h and kzg_serve.h: the SRS
// and context handles, and the batched commit pipeline shared by
// kzg_commit_batch and the serving loop.
#pragma once

#include <map>
#include <memory>
#include <vector>

#include "common.h"
#include "msm.h"

struct kzg_srs {
    int curve = 0;
    size_t n = 0;
    kz::DevBuf pts;   // Montgomery affine records
    kz::DevBuf inf;   // identity flags
    kz::DevBuf wire;  // wire points (kept for export)
    bool g1 = false;  // every point lies in G1 (generated, or checked at creation)
    // Optional window tables of the points (kzg_srs_precompute), one after
    // another in `tables` (records of affine_words each): table k has window
    // bits table_c[k] and starts at record table_base[k]; its row 0 is the
    // points themselves.
    kz::DevBuf tables;
    std::vector<int> table_c;
    std::vector<uint32_t> table_base;
};

struct kzg_ctx {
    hipStream_t stream = nullptr;
    kz::MsmWorkspace ws;
    kz::DevBuf coeffs, tmp;
};

namespace kz {

// MSM of the first n SRS points with device scalars: through the SRS window
// table whose window bits are nearest to msm_table_window_bits(n) (or equal
// window_bits when given), else the prepared points (GLV when they are known
// to lie in G1).
void srs_msm(const kzg_srs* srs, const uint32_t* d_scalars, size_t n, int window_bits, MsmWorkspace& ws,
             hipStream_t s, uint8_t* d_out);

// Index of the SRS table whose window bits are nearest to c (ties: the wider),
// -1 if the SRS has none within max_gap bits.
int srs_table_near(const kzg_srs* srs, int c, int max_gap);

// value < r for a 32-byte LE element of the curve's scalar field
bool canonical_fr(int curve, const uint8_t* le32);
// log2(n) if n is a power of two, else -1
int log2_exact(size_t n);

// Host memory from hipHostMalloc (page-locked), grow-only.
struct PinnedBuf {
    void* ptr = nullptr;
    size_t cap = 0;
    PinnedBuf() = default;
    PinnedBuf(const PinnedBuf&) = delete;
    PinnedBuf& operator=(const PinnedBuf&) = delete;
    ~PinnedBuf() {
        if (ptr) (void)hipHostFree(ptr);
    }
    void* ensure(size_t bytes) {
        if (bytes > cap) {
            if (ptr) KZ_HIP(hipHostFree(ptr));
            ptr = nullptr;
            cap = 0;
            KZ_HIP(hipHostMalloc(&ptr, bytes < 4096 ? 4096 : bytes, hipHostMallocDefault));
            cap = bytes < 4096 ? 4096 : bytes;
        }
        return ptr;
    }
    template <class T>
    T* as() const {
        return static_cast<T*>(ptr);
    }
};

// A lane's cached batch plan for one set of job sizes and, once those sizes
// come back, the captured launch sequence of the whole commit pipeline (H2D,
// inverse NTTs, batched MSM, D2H) over the lane's buffers, replayed as one
// graph launch while the buffers stay where they are.
struct LanePlan {
    MsmBatchPlan plan;
    uint32_t uses = 0;
    bool no_graph = false;  // capture failed once: launch directly
    // every device and host address the captured graph uses: the lane's
    // buffers, the SRS arrays and the lane's MSM workspace (which a larger
    // plan of the same lane may reallocate)
    static constexpr int kAddrs = 8 + MsmWorkspace::kBuffers;
    const void* bufs[kAddrs] = {};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    LanePlan() = default;
    LanePlan(const LanePlan&) = delete;
    LanePlan& operator=(const LanePlan&) = delete;
    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
    ~LanePlan() { drop_graph(); }
};

// One stream with everything a batch of commitments needs: device buffers
// for evaluations / coefficients / scratch, the MSM workspace, page-locked
// staging both ways, cached plans and graphs (keyed by the job sizes) and a
// completion event.  A lane runs one batch at a time.
struct CommitLane {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    MsmWorkspace ws;
    DevBuf evals, coeffs, tmp, out, err;
    PinnedBuf pin_in, pin_out;
    std::map<std::pair<int, std::vector<uint32_t>>, std::unique_ptr<LanePlan>> plans;
    uint64_t graph_launches = 0, direct_launches = 0;

    explicit CommitLane(bool own_stream);
    ~CommitLane();
    CommitLane(const CommitLane&) = delete;
    CommitLane& operator=(const CommitLane&) = delete;
    bool owns_stream = false;
};

// Launches commitments to `count` polynomials of one curve against `srs` on
// lane.stream: evaluations (32-byte LE, natural order; sizes n[i], powers of
// two, sorted ascending) already staged back to back in lane.pin_in.  H2D,
// one inverse-NTT launch sequence per distinct size, one batched MSM over the
// SRS prefix, D2H of the records (wire + compressed) and the error flag into
// lane.pin_out: record i at pin_out + i * msm_out_bytes, the flag (int) after
// the last record.  A set of job sizes launches directly at first; once it
// recurs (4 uses by default), the sequence is captured into a HIP graph (once
// per buffer layout) and replayed.  Records the lane's done event.  Nothing
// synchronises.
void commit_batch_launch(CommitLane& lane, const kzg_srs* srs, const std::vector<uint32_t>& n);
// Bytes of lane.pin_in / lane.pin_out a batch of these sizes needs.
size_t commit_batch_in_bytes(const std::vector<uint32_t>& n);
size_t commit_batch_out_bytes(int curve, size_t count);

}  // namespace kz
