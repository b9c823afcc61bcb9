// Host side of the Pippenger MSM: workspace sizing, the launch sequence and
// curve dispatch.  Kernels are in msm_impl.h / msm_digits.hip / scan.hip.
#include <atomic>
#include <vector>

#include "msm_internal.h"
#include "msm_stages.h"
#include "scan.h"

namespace kz {

// Window bits from n: about 32 points per bucket on average (2^(c-1) ~ n/32),
// clamped to the instantiated range.
int msm_default_window_bits(size_t n) {
    int lg = 0;
    while (((size_t)1 << (lg + 1)) <= n) ++lg;
    int c = lg - 4;
    if (c < 5) c = 5;
    if (c > 16) c = 16;
    return c;
}

// ---------------------------------------------------------------- profiler
// Opt-in per-stage timing: HIP events recorded on the MSM's own stream around
// each stage, summed when collected.  Off by default (no events recorded).
namespace {

std::atomic<int> g_profiling{0};
std::mutex g_prof_mu;
struct StageEvents {
    int stage;
    hipEvent_t a, b;
};
std::vector<StageEvents> g_prof_events;

struct StageTimer {
    hipStream_t s;
    int stage;
    hipEvent_t a = nullptr;
    StageTimer(hipStream_t s_, int stage_) : s(s_), stage(stage_) {
        if (!g_profiling.load()) return;
        KZ_HIP(hipEventCreate(&a));
        KZ_HIP(hipEventRecord(a, s));
    }
    ~StageTimer() {
        if (!a) return;
        hipEvent_t b;
        if (hipEventCreate(&b) != hipSuccess) return;
        (void)hipEventRecord(b, s);
        std::lock_guard<std::mutex> g(g_prof_mu);
        g_prof_events.push_back({stage, a, b});
    }
};

}  // namespace

void msm_set_profiling(int on) { g_profiling.store(on); }

int msm_collect_profile(double* sums_ms, int* counts, int nstages) {
    std::lock_guard<std::mutex> g(g_prof_mu);
    for (int i = 0; i < nstages; ++i) {
        sums_ms[i] = 0;
        counts[i] = 0;
    }
    for (auto& ev : g_prof_events) {
        KZ_HIP(hipEventSynchronize(ev.b));
        float ms = 0;
        KZ_HIP(hipEventElapsedTime(&ms, ev.a, ev.b));
        if (ev.stage < nstages) {
            sums_ms[ev.stage] += ms;
            counts[ev.stage] += 1;
        }
        (void)hipEventDestroy(ev.a);
        (void)hipEventDestroy(ev.b);
    }
    g_prof_events.clear();
    return 0;
}

namespace {

constexpr uint32_t kSegBuckets = 16;  // must match kSeg in msm_impl.h

struct Stages {
    decltype(&launch_accumulate_bls) accumulate;
    decltype(&launch_fold_bls) fold;
    decltype(&launch_reduce_seg_bls) reduce_seg;
    decltype(&launch_reduce_window_bls) reduce_window;
    decltype(&launch_combine_bls) combine;
    int xyzz_words;
};

const Stages kBls = {launch_accumulate_bls, launch_fold_bls, launch_reduce_seg_bls, launch_reduce_window_bls, launch_combine_bls, 56};
const Stages kBn = {launch_accumulate_bn, launch_fold_bn, launch_reduce_seg_bn, launch_reduce_window_bn, launch_combine_bn, 36};

void run_pipeline(const Stages& st, const uint32_t* d_pts, const uint8_t* d_inf, const uint32_t* d_scalars,
                  size_t n, const MsmParams& prm, MsmWorkspace& ws, hipStream_t s, uint8_t* d_out) {
    const int c = prm.c ? prm.c : msm_default_window_bits(n);
    if (c < 5 || c > 18) throw std::invalid_argument("window bits out of range [5, 18]");
    const int W = msm_num_windows(c);
    const uint32_t B = 1u << (c - 1);
    const uint32_t NB = (uint32_t)W * B;
    const uint32_t chunk = (uint32_t)(prm.chunk > 0 ? prm.chunk : 32);
    const size_t max_entries = n * (size_t)W;
    if (max_entries >= 0x80000000ull) throw std::invalid_argument("MSM too large for 31-bit entry indices");
    const uint32_t max_chunks = ceil_div(max_entries, chunk);
    const size_t pw = (size_t)st.xyzz_words * sizeof(uint32_t);  // XYZZ record bytes
    const uint32_t t1 = B / kSegBuckets;
    const uint32_t nseg = (uint32_t)W * t1;

    ws.counts.ensure(NB * sizeof(uint32_t));
    ws.offsets.ensure((NB + 1) * sizeof(uint32_t));
    ws.cursor.ensure(NB * sizeof(uint32_t));
    ws.block_sums.ensure(scan_block_count(NB) * sizeof(uint32_t));
    ws.total.ensure(sizeof(uint32_t));
    ws.entries.ensure((max_entries + 1) * sizeof(uint32_t));
    ws.bsum.ensure(NB * pw);
    ws.carry.ensure((size_t)(max_chunks + 1) * pw);
    ws.run_bucket.ensure((size_t)(max_chunks + 1) * sizeof(uint32_t));
    ws.segA.ensure((size_t)nseg * pw);
    ws.segB.ensure((size_t)nseg * pw);
    ws.winres.ensure((size_t)W * pw);

    uint32_t* counts = ws.counts.as<uint32_t>();
    uint32_t* offsets = ws.offsets.as<uint32_t>();
    {
        StageTimer t(s, kStageSort);
        KZ_HIP(hipMemsetAsync(counts, 0, NB * sizeof(uint32_t), s));
        digits_dispatch(c, false, d_scalars, d_inf, (uint32_t)n, counts, nullptr, s);
        device_exclusive_scan(counts, offsets, ws.cursor.as<uint32_t>(), NB, ws.block_sums.as<uint32_t>(),
                              ws.total.as<uint32_t>(), s);
        digits_dispatch(c, true, d_scalars, d_inf, (uint32_t)n, ws.cursor.as<uint32_t>(), ws.entries.as<uint32_t>(),
                        s);
    }
    {
        StageTimer t(s, kStageAccumulate);
        st.accumulate(ceil_div(max_chunks, 64), s, ws.entries.as<uint32_t>(), offsets, NB, chunk, d_pts,
                      ws.bsum.as<uint32_t>(), ws.carry.as<uint32_t>());
    }
    // A bucket holds at most n entries (one per point), so its carry run is at
    // most ceil(n / chunk) long: that many tree levels always suffice.
    uint32_t levels = 0;
    while (((size_t)1 << levels) < (n + chunk - 1) / chunk + 1) ++levels;
    {
        StageTimer t(s, kStageReduce);
        st.fold(s, offsets, NB, chunk, max_chunks, levels, ws.run_bucket.as<uint32_t>(), ws.carry.as<uint32_t>());
        st.reduce_seg(ceil_div(nseg, 64), s, offsets, chunk, ws.bsum.as<uint32_t>(), ws.carry.as<uint32_t>(), nseg,
                      ws.segA.as<uint32_t>(), ws.segB.as<uint32_t>());
        st.reduce_window((unsigned)W, s, ws.segA.as<uint32_t>(), ws.segB.as<uint32_t>(), t1,
                         ws.winres.as<uint32_t>());
    }
    {
        StageTimer t(s, kStageCombine);
        st.combine(s, ws.winres.as<uint32_t>(), W, c, d_out);
    }
    KZ_HIP(hipGetLastError());
}

}  // namespace

void msm_device(int curve, const uint32_t* d_pts, const uint8_t* d_inf, const uint32_t* d_scalars, size_t n,
                const MsmParams& prm, MsmWorkspace& ws, hipStream_t s, uint8_t* d_out) {
    run_pipeline(curve == 2 ? kBls : kBn, d_pts, d_inf, d_scalars, n, prm, ws, s, d_out);
}

void points_from_wire(int curve, const uint8_t* d_wire, uint32_t* d_pts, uint8_t* d_inf, size_t n,
                      hipStream_t s) {
    if (n == 0) return;
    if (curve == 2) launch_points_from_wire_bls(d_wire, d_pts, d_inf, (uint32_t)n, s);
    else launch_points_from_wire_bn(d_wire, d_pts, d_inf, (uint32_t)n, s);
    KZ_HIP(hipGetLastError());
}

void compress_on_device(int curve, const uint8_t* d_wire, uint8_t* d_out, hipStream_t s) {
    if (curve == 2) launch_compress_bls(d_wire, d_out, s);
    else launch_compress_bn(d_wire, d_out, s);
    KZ_HIP(hipGetLastError());
}

}  // namespace kz
