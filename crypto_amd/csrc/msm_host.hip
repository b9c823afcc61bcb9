// Host side of the Pippenger MSM: workspace sizing, the launch sequence, curve
// dispatch, batches of MSMs and the window-size tables.  Kernels are in
// msm_impl.h / msm_digits.hip / scan.hip.
#include <atomic>
#include <vector>

#include "glv.h"
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

// ------------------------------------------------------------ tuning tables
// Per (curve, mode, floor(log2 n)): window bits and first-level reduction
// segment, 0 = the built-in rule.  Filled by the autotuner or a cache file.
namespace {

std::mutex g_table_mu;
struct TuneCell {
    int c = 0, seg = 0;
};
TuneCell g_table[3][kMsmModes][32];

int floor_log2(size_t n) {
    int lg = 0;
    while (((size_t)1 << (lg + 1)) <= n) ++lg;
    return lg;
}

void check_mode(int curve, int mode) {
    if (!msm_tunable(curve, mode)) throw std::invalid_argument("bad curve or mode");
}

}  // namespace

int msm_window_bits(int curve, size_t n, int mode) {
    const int lg = floor_log2(n);
    int c = 0;
    if (msm_tunable(curve, mode) && lg < 32) {
        std::lock_guard<std::mutex> g(g_table_mu);
        c = g_table[curve][mode][lg].c;
    }
    if (c == kMsmGlvOff && mode == kMsmGlv) return msm_window_bits(curve, n, kMsmPlain);
    if (c) return c;
    // GLV runs 2n points
    return msm_default_window_bits(mode == kMsmGlv ? 2 * n : n);
}

bool msm_glv_enabled(int curve, size_t n) {
    const int lg = floor_log2(n);
    if (curve < 0 || curve > 2 || lg >= 32) return true;
    std::lock_guard<std::mutex> g(g_table_mu);
    return g_table[curve][kMsmGlv][lg].c != kMsmGlvOff;
}

int msm_tuned_segment(int curve, size_t n, int mode) {
    const int lg = floor_log2(n);
    if (!msm_tunable(curve, mode) || lg >= 32) return 0;
    std::lock_guard<std::mutex> g(g_table_mu);
    return g_table[curve][mode][lg].seg;
}

void msm_set_window_table(int curve, int mode, const int* c_by_log, int count) {
    check_mode(curve, mode);
    if (count < 0 || count > 32) throw std::invalid_argument("bad window table");
    for (int k = 0; k < count; ++k)
        if (c_by_log[k] != 0 && (c_by_log[k] < 5 || c_by_log[k] > 18) &&
            !(mode == kMsmGlv && c_by_log[k] == kMsmGlvOff))
            throw std::invalid_argument("window bits out of range [5, 18]");
    std::lock_guard<std::mutex> g(g_table_mu);
    for (int k = 0; k < 32; ++k) g_table[curve][mode][k].c = k < count ? c_by_log[k] : 0;
}

void msm_set_tuned(int curve, int mode, int lg, int c, int seg) {
    check_mode(curve, mode);
    if (lg < 0 || lg >= 32) throw std::invalid_argument("size class out of range");
    if (c != 0 && (c < 5 || c > 18) && !(mode == kMsmGlv && c == kMsmGlvOff))
        throw std::invalid_argument("window bits out of range [5, 18]");
    if (seg != 0 && (seg > 16 || (seg & (seg - 1)) != 0))
        throw std::invalid_argument("segment must be 0 or a power of two in [1, 16]");
    std::lock_guard<std::mutex> g(g_table_mu);
    g_table[curve][mode][lg] = TuneCell{c, seg};
}

void msm_get_tuned(int curve, int mode, int lg, int& c, int& seg) {
    check_mode(curve, mode);
    if (lg < 0 || lg >= 32) throw std::invalid_argument("size class out of range");
    std::lock_guard<std::mutex> g(g_table_mu);
    c = g_table[curve][mode][lg].c;
    seg = g_table[curve][mode][lg].seg;
}

// ---------------------------------------------------------------- profiler
// Opt-in per-stage timing: HIP events recorded on the MSM's own stream around
// each stage, summed when collected.  Off by default (no events recorded).
namespace {

std::atomic<int> g_profiling{0};
thread_local int g_timers_off = 0;  // NoStageTimers nesting on this thread
// First-level reduction segment (buckets per running-sum chain): the chain is
// 3 seg point additions long, so short segments cut latency; 0 = built-in rule.
std::atomic<int> g_seg_override{0};
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
        if (!g_profiling.load() || g_timers_off) return;
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        if (s && hipStreamIsCapturing(s, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone) return;
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

NoStageTimers::NoStageTimers() { ++g_timers_off; }
NoStageTimers::~NoStageTimers() { --g_timers_off; }

void msm_set_reduce_segment(int seg) {
    if (seg != 0 && (seg < 1 || seg > 16 || (seg & (seg - 1)) != 0))
        throw std::invalid_argument("segment must be 0 or a power of two in [1, 16]");
    g_seg_override.store(seg);
}

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

// ---------------------------------------------------------------- pipeline
namespace {

// Built-in rule: 16-bucket segments while that still leaves >= 2^14 segments
// (4 lanes each, so >= 1024 waves over 256 CUs); below that, halve down to 4,
// trading a short chain for more waves.  Measured on BLS12-381 / BN254 MSMs
// from 2^12 to 2^20 points: seg 4 cuts the reduction by 15-45 % at n <= 2^18,
// seg 16 is best at 2^20 (c = 16).  min_buckets_per_window >= 16 (c >= 5).
// A tuned segment (nonzero `tuned`) wins over the rule; the global override
// wins over both.
uint32_t reduce_segment(uint64_t total_buckets, uint32_t min_buckets_per_window, int tuned = 0) {
    uint32_t seg = (uint32_t)g_seg_override.load();
    if (!seg) seg = (uint32_t)tuned;
    if (!seg) {
        seg = 16;
        while (seg > 4 && total_buckets / seg < (1u << 14)) seg >>= 1;
    }
    while (seg > min_buckets_per_window) seg >>= 1;
    return seg;
}

// Carry-fold scratch in ws.run_bucket: 3 control words (padded to 4), the
// list of long-run chunks (4 words, 16-byte aligned, per entry), then the
// long runs {head, len} (2 words each; a long run has more than
// kFoldShortRun chunks, so there are fewer than max_chunks / kFoldShortRun).
size_t fold_words(uint32_t max_chunks) {
    return 4 + 4 * (size_t)max_chunks + 2 * ((size_t)max_chunks / kFoldShortRun + 1);
}
uint32_t* fold_ctl(MsmWorkspace& ws) { return ws.run_bucket.as<uint32_t>(); }
uint32_t* fold_list(MsmWorkspace& ws) { return ws.run_bucket.as<uint32_t>() + 4; }
uint32_t* fold_runs(MsmWorkspace& ws, uint32_t max_chunks) {
    return ws.run_bucket.as<uint32_t>() + 4 + 4 * (size_t)max_chunks;
}

uint32_t log2u(uint32_t v) {
    uint32_t l = 0;
    while ((1u << l) < v) ++l;
    return l;
}

struct Stages {
    decltype(&launch_accumulate_bls) accumulate;
    decltype(&launch_fold_bls) fold;
    decltype(&launch_reduce_seg_bls) reduce_seg;
    decltype(&launch_reduce_level_bls) reduce_level;
    decltype(&launch_combine_bls) combine;
    decltype(&launch_table_rows_bls) table_rows;
    int xyzz_words;
};

const Stages kBls = {launch_accumulate_bls, launch_fold_bls,    launch_reduce_seg_bls,
                     launch_reduce_level_bls, launch_combine_bls, launch_table_rows_bls, 56};
const Stages kBn = {launch_accumulate_bn, launch_fold_bn,    launch_reduce_seg_bn,
                    launch_reduce_level_bn, launch_combine_bn, launch_table_rows_bn, 36};

const Stages kSecp = {launch_accumulate_secp, launch_fold_secp,    launch_reduce_seg_secp,
                      launch_reduce_level_secp, launch_combine_secp, launch_table_rows_secp, 40};

// Banderwagon: extended Edwards points (36 words); no window tables.
const Stages kBw = {launch_accumulate_bw, launch_fold_bw,    launch_reduce_seg_bw,
                    launch_reduce_level_bw, launch_combine_bw, nullptr, 36};

const Stages& stages_of(int curve) {
    return curve == 2 ? kBls : curve == 1 ? kBn : curve == 3 ? kBw : kSecp;
}

// Sizes of one pipeline run, single MSM or batch.
struct Shape {
    uint32_t nb;        // global buckets
    uint32_t nwin;      // global windows
    uint32_t t1;        // segments per window (single MSM), 0 = per-window tables
    uint32_t seg;       // buckets per first-level segment
    uint32_t nseg;      // first-level segments
    uint32_t chunk;     // sorted entries per accumulation thread
    uint32_t levels;    // carry-fold tree levels
    size_t max_entries;
    size_t level_out;   // largest output count of a reduction level
    uint32_t max_chunks() const { return ceil_div(max_entries, chunk); }
};

// Sorted entries per accumulation thread: 32, doubled (up to 256) while a
// bucket would still average at least 4 chunks -- fewer carries to store and
// fold -- and at least 2^18 threads remain.  GLV MSMs of 16-bit windows then
// take 32 at 2^20 points, 64 at 2^21 and 128 at 2^22 (buckets of 64, 128, 256
// entries); window tables of large point sets get up to 256.  Measured on one
// box, GLV 2^21 / 2^22, chunk 32 against the rule's: BLS12-381 7.63 / 14.32
// against 7.51 / 13.96 ms, BN254 4.21 / 8.50 against 4.17 / 8.40 ms; at 2^20
// chunk 64 was no faster (BN254 2.27 against 2.32 ms).
uint32_t auto_chunk(size_t max_entries, uint32_t nb) {
    const size_t per_bucket = max_entries / (nb ? nb : 1);
    uint32_t chunk = 32;
    while (chunk < 256 && per_bucket / chunk >= 4 && max_entries / (2 * (size_t)chunk) >= ((size_t)1 << 18)) chunk *= 2;
    return chunk;
}

// A bucket holds at most n entries (one per point), so its carry run is at
// most ceil(n / chunk) + 1 long: log2 of that many tree levels suffice.
uint32_t fold_levels(size_t max_n, uint32_t chunk) {
    uint32_t levels = 0;
    while (((size_t)1 << levels) < (max_n + chunk - 1) / chunk + 1) ++levels;
    return levels;
}

void reserve(const Stages& st, const Shape& sh, MsmWorkspace& ws) {
    if (sh.max_entries >= 0x80000000ull) throw std::invalid_argument("MSM too large for 31-bit entry indices");
    const size_t pw = (size_t)st.xyzz_words * sizeof(uint32_t);  // XYZZ record bytes
    const uint32_t max_chunks = sh.max_chunks();
    ws.counts.ensure((size_t)sh.nb * sizeof(uint32_t));
    ws.offsets.ensure(((size_t)sh.nb + 1) * sizeof(uint32_t));
    ws.cursor.ensure((size_t)sh.nb * sizeof(uint32_t));
    ws.block_sums.ensure((size_t)scan_block_count(sh.nb) * sizeof(uint32_t));
    ws.total.ensure(sizeof(uint32_t));
    ws.entries.ensure((sh.max_entries + 1) * sizeof(uint32_t));
    ws.bsum.ensure((size_t)sh.nb * pw);
    ws.carry.ensure((size_t)(max_chunks + 1) * pw);
    ws.run_bucket.ensure(fold_words(max_chunks) * sizeof(uint32_t));
    ws.segA.ensure((size_t)sh.nseg * pw);
    ws.segB.ensure((size_t)sh.nseg * pw);
    ws.winres.ensure((size_t)sh.nwin * pw);
    const size_t lv = (sh.level_out > sh.nwin ? sh.level_out : sh.nwin) * pw;
    for (int i = 0; i < 2; ++i) {
        ws.lvA[i].ensure(lv);
        ws.lvB[i].ensure(lv);
    }
}

uint32_t parts_of(uint32_t n_in) { return (n_in + kReduceFanIn - 1) / kReduceFanIn; }

// Global-atomic sort: clear the histogram, count, scan, scatter.
// `digits(scatter, counts, entries)` launches the count or scatter pass.
template <class Digits>
void atomic_sort(const Shape& sh, Digits&& digits, MsmWorkspace& ws, hipStream_t s) {
    uint32_t* counts = ws.counts.as<uint32_t>();
    device_fill_u32(counts, 0, sh.nb, s);
    digits(false, counts, nullptr);
    device_exclusive_scan(counts, ws.offsets.as<uint32_t>(), ws.cursor.as<uint32_t>(), sh.nb,
                          ws.block_sums.as<uint32_t>(), ws.total.as<uint32_t>(), s);
    digits(true, ws.cursor.as<uint32_t>(), ws.entries.as<uint32_t>());
}

// Scope of one MSM's use of a workspace: its stream first waits for the event
// the previous user recorded (when that was another stream), and records the
// event again at the end.  Skipped while the stream is being captured into a
// graph: the order of graph replays is the caller's.
struct WorkspaceUse {
    MsmWorkspace& ws;
    hipStream_t s;
    bool on = true;
    WorkspaceUse(MsmWorkspace& w, hipStream_t st) : ws(w), s(st) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        if (hipStreamIsCapturing(s, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone) on = false;
        if (on && ws.last_use && ws.last_stream != s) KZ_HIP(hipStreamWaitEvent(s, ws.last_use, 0));
    }
    ~WorkspaceUse() {
        if (!on) return;
        if (!ws.last_use && hipEventCreateWithFlags(&ws.last_use, hipEventDisableTiming) != hipSuccess) {
            ws.last_use = nullptr;
            return;
        }
        if (hipEventRecord(ws.last_use, s) == hipSuccess) ws.last_stream = s;
    }
};

// `sort()` fills ws.offsets and ws.entries.
template <class Sort>
void run_stages(const Stages& st, const Shape& sh, Sort&& sort, const MsmBatchPlan* plan,
                const MsmJob* d_jobs, const MsmJob& one, uint32_t njobs, const uint32_t* d_pts, MsmWorkspace& ws,
                hipStream_t s, uint8_t* d_out) {
    uint32_t* offsets = ws.offsets.as<uint32_t>();
    const uint32_t max_chunks = sh.max_chunks();
    {
        StageTimer t(s, kStageSort);
        sort();
    }
    {
        StageTimer t(s, kStageAccumulate);
        st.accumulate(ceil_div(max_chunks, 64), s, ws.entries.as<uint32_t>(), offsets, sh.nb, sh.chunk, d_pts,
                      ws.bsum.as<uint32_t>(), ws.carry.as<uint32_t>());
    }
    {
        StageTimer t(s, kStageReduce);
        st.fold(s, offsets, sh.nb, sh.chunk, max_chunks, sh.levels, fold_list(ws), fold_runs(ws, max_chunks),
                fold_ctl(ws), ws.carry.as<uint32_t>());
        st.reduce_seg(ceil_div((size_t)sh.nseg * 4, 256), s, offsets, sh.chunk, ws.bsum.as<uint32_t>(),
                      ws.carry.as<uint32_t>(), sh.nseg, sh.seg, ws.segA.as<uint32_t>(), ws.segB.as<uint32_t>());
        // levels: segment pairs carry weight m = seg; each level multiplies it
        // by kReduceFanIn = 2^7
        const uint32_t* inA = ws.segA.as<uint32_t>();
        const uint32_t* inB = ws.segB.as<uint32_t>();
        uint32_t lm = log2u(sh.seg), n_in = sh.t1;
        const size_t nlev = plan ? plan->level_grid.size() : 0;
        for (size_t lv = 0;; ++lv) {
            const uint32_t parts = plan ? 0 : parts_of(n_in);
            const bool final = plan ? lv + 1 == nlev : parts == 1;
            uint32_t* oA = ws.lvA[lv & 1].as<uint32_t>();
            uint32_t* oB = final ? ws.winres.as<uint32_t>() : ws.lvB[lv & 1].as<uint32_t>();
            if (plan)
                st.reduce_level(plan->level_grid[lv], s, inA, inB, plan->d_levels.as<RedLevel>() + plan->level_off[lv],
                                sh.nwin, 0, 0, lm, oA, oB);
            else
                st.reduce_level(sh.nwin * parts, s, inA, inB, nullptr, 0, n_in, parts, lm, oA, oB);
            if (final) break;
            inA = oA;
            inB = oB;
            n_in = parts;
            lm += 7;
        }
    }
    {
        StageTimer t(s, kStageCombine);
        st.combine(s, ws.winres.as<uint32_t>(), d_jobs, one, njobs, d_out);
    }
    KZ_HIP(hipGetLastError());
}

}  // namespace

// One pipeline over at most kMsmSlice points (point indices and bucket
// offsets of a pipeline are 32-bit).
static void msm_device_one(int curve, const uint32_t* d_pts, const uint8_t* d_inf, const uint32_t* d_scalars,
                           size_t n, const MsmParams& prm, MsmWorkspace& ws, hipStream_t s, uint8_t* d_out) {
    const Stages& st = stages_of(curve);
    WorkspaceUse use(ws, s);
    // GLV (points in G1): 2n points with scalar magnitudes below 2^127
    // (BN254, BLS12-381) or 2^128 (secp256k1)
    int bits = 256;
    size_t npts = n;
    int c = prm.c;
    const bool table = prm.table_stride != 0;
    if (table) {
        if (c < 5 || c > 18) throw std::invalid_argument("a window-table MSM needs the table's window bits");
        if (n > prm.table_stride) throw std::invalid_argument("more scalars than table points");
        if ((uint64_t)msm_table_rows(c) * prm.table_stride >= 0x80000000ull)
            throw std::invalid_argument("window table beyond 2^31 points");
    }
    if (prm.g1 && !table && n <= (1u << 29) && (c || msm_glv_enabled(curve, n))) {
        const int cg = c ? c : msm_window_bits(curve, n, kMsmGlv);
        if (lds_sort_supported(cg, 2 * n)) {
            // |k1|, |k2| < 2^127 for BN254 and BLS12-381 (glv.h): 127-bit
            // windows.  The top window then holds t < c real bits (c never
            // divides the prime 127), so its digit plus an incoming carry is at
            // most 2^t <= 2^(c-1) and never carries out: no carry window (8
            // windows at c = 16 instead of 9).  secp256k1's halves reach
            // 2^127.35: 128-bit windows, with the carry window when c divides 128.
            bits = curve == 0 ? 128 : 127;
            npts = 2 * n;
            c = cg;
            const size_t pw = (size_t)affine_words(curve) * 4;
            uint32_t* pts2 = static_cast<uint32_t*>(ws.glv_pts.ensure(2 * n * pw));
            uint8_t* inf2 = static_cast<uint8_t*>(ws.glv_inf.ensure(2 * n));
            uint32_t* sc2 = static_cast<uint32_t*>(ws.glv_sc.ensure(2 * n * 16));  // 16-byte magnitudes
            glv_expand(curve, d_pts, d_inf, d_scalars, n, pts2, inf2, sc2, s);
            d_pts = pts2;
            d_inf = inf2;
            d_scalars = sc2;
        }
    }
    if (!c) c = msm_window_bits(curve, n, kMsmPlain);
    if (c < 5 || c > 18) throw std::invalid_argument("window bits out of range [5, 18]");
    // the tuned segment belongs to the tuned window width
    const int tuned_seg = prm.seg ? prm.seg : prm.c ? 0 : msm_tuned_segment(curve, n, bits != 256 ? kMsmGlv : kMsmPlain);
    const int W = msm_num_windows(c, bits);
    const uint32_t B = 1u << (c - 1);
    const uint32_t bw = table ? 1u : (uint32_t)W;  // bucket windows (a table shares one)
    Shape sh;
    sh.nb = bw * B;
    sh.nwin = bw;
    sh.seg = reduce_segment(sh.nb, B, table ? prm.seg : tuned_seg);
    sh.t1 = B / sh.seg;
    sh.nseg = bw * sh.t1;
    sh.max_entries = npts * (size_t)W;
    sh.chunk = prm.chunk > 0 ? (uint32_t)prm.chunk : auto_chunk(sh.max_entries, sh.nb);
    sh.levels = fold_levels(table ? npts * (size_t)W : npts, sh.chunk);
    sh.level_out = (size_t)bw * parts_of(sh.t1);
    reserve(st, sh, ws);
    MsmJob one{};
    one.n = (uint32_t)npts;
    one.c = (uint32_t)c;
    one.nwin = bw;
    const bool lds = lds_sort_supported(c, npts);
    if (lds) {
        const size_t words = lds_sort_words(c, bits, npts);
        ws.sort_cnt.ensure(words * sizeof(uint32_t));
        ws.sort_pre.ensure(words * sizeof(uint32_t));
    }
    auto sort = [&] {
        if (lds) {
            lds_sort_dispatch(c, bits, d_scalars, d_inf, (uint32_t)npts, prm.table_stride, ws.sort_cnt.as<uint32_t>(),
                              ws.sort_pre.as<uint32_t>(), ws.counts.as<uint32_t>(), ws.offsets.as<uint32_t>(),
                              ws.block_sums.as<uint32_t>(), ws.total.as<uint32_t>(), ws.entries.as<uint32_t>(), s);
        } else {
            atomic_sort(sh, [&](bool scatter, uint32_t* cc, uint32_t* entries) {
                digits_dispatch(c, scatter, d_scalars, d_inf, (uint32_t)npts, prm.table_stride, cc, entries, s);
            }, ws, s);
        }
    };
    run_stages(st, sh, sort, nullptr, nullptr, one, 1, d_pts, ws, s, d_out);
}

// Record j of the slice outputs (stride bytes apart) -> wire point j; unit
// scalars (1, 0, ..., 0) for the sum of the slices.
__global__ void k_slice_sum_inputs(const uint8_t* __restrict__ outs, uint32_t stride, uint32_t pb, uint32_t nslices,
                                   uint8_t* __restrict__ wire, uint32_t* __restrict__ unit) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nslices * pb) wire[t] = outs[(size_t)(t / pb) * stride + t % pb];
    if (t < nslices * 8) unit[t] = (t % 8 == 0) ? 1u : 0u;
}

// Larger MSMs run as slices of kMsmSlice points through the same workspace,
// one after another on the stream; the slice results (wire points) are then
// prepared like any input and summed by one more MSM with unit scalars, so
// the output is the exact sum, byte for byte what one pipeline would give.
void msm_device(int curve, const uint32_t* d_pts, const uint8_t* d_inf, const uint32_t* d_scalars, size_t n,
                const MsmParams& prm, MsmWorkspace& ws, hipStream_t s, uint8_t* d_out) {
    if (prm.table_stride || n <= kMsmSlice) {
        msm_device_one(curve, d_pts, d_inf, d_scalars, n, prm, ws, s, d_out);
        return;
    }
    const size_t J = (n + kMsmSlice - 1) / kMsmSlice;
    if (J > 4096) throw std::invalid_argument("MSM beyond 2^37 points");
    const size_t ob = msm_out_bytes(curve), pb = point_wire_bytes(curve);
    const size_t aw = (size_t)affine_words(curve);
    // slice outputs | wire points | prepared points | flags | unit scalars | error word
    const size_t o_wire = J * ob, o_pts = (o_wire + J * pb + 255) / 256 * 256, o_inf = o_pts + J * aw * 4;
    const size_t o_unit = (o_inf + J + 255) / 256 * 256, o_err = o_unit + J * 32;
    uint8_t* b = static_cast<uint8_t*>(ws.slices.ensure(o_err + 4));
    for (size_t j = 0; j < J; ++j) {
        const size_t first = j * kMsmSlice, len = std::min(kMsmSlice, n - first);
        msm_device_one(curve, d_pts + first * aw, d_inf ? d_inf + first : nullptr, d_scalars + first * 8, len, prm,
                       ws, s, b + j * ob);
    }
    uint8_t* wire = b + o_wire;
    uint32_t* pts = reinterpret_cast<uint32_t*>(b + o_pts);
    uint8_t* inf = b + o_inf;
    uint32_t* unit = reinterpret_cast<uint32_t*>(b + o_unit);
    const uint32_t threads = (uint32_t)std::max(J * pb, J * 8);
    hipLaunchKernelGGL(k_slice_sum_inputs, dim3(ceil_div(threads, 256u)), dim3(256), 0, s, b, (uint32_t)ob,
                       (uint32_t)pb, (uint32_t)J, wire, unit);
    KZ_HIP(hipGetLastError());
    if (curve == 3) {  // the validating reader; slice sums of valid points pass it
        uint32_t* err = reinterpret_cast<uint32_t*>(b + o_err);
        device_fill_u32(err, 0, 1, s);
        bw_points_from_wire(wire, pts, inf, (uint32_t)J, err, s);
    } else {
        points_from_wire(curve, wire, pts, inf, J, s);
    }
    msm_device_one(curve, pts, inf, unit, J, MsmParams{}, ws, s, d_out);
}

int msm_table_window_bits(size_t n) {
    for (int c = 16; c > 5; --c)
        if ((uint64_t)msm_table_rows(c) * n >= (32ull << (c - 1))) return c;
    return 5;
}

void points_table_build(int curve, const uint32_t* d_pts, const uint8_t* d_inf, size_t n, int c, uint32_t* d_table,
                        hipStream_t s) {
    if (c < 5 || c > 18) throw std::invalid_argument("window bits out of range [5, 18]");
    if (n == 0) return;
    if ((uint64_t)msm_table_rows(c) * n >= 0x80000000ull) throw std::invalid_argument("window table beyond 2^31 points");
    stages_of(curve).table_rows(d_pts, d_inf, (uint32_t)n, (uint32_t)c, (uint32_t)msm_table_rows(c), d_table, s);
    KZ_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ batches
namespace {

Shape batch_shape(const MsmBatchPlan& p) {
    Shape sh;
    sh.nb = p.nb;
    sh.nwin = p.nwin;
    sh.t1 = 0;
    sh.seg = p.seg;
    sh.nseg = p.nseg;
    sh.chunk = p.chunk;
    sh.levels = p.levels;
    sh.max_entries = p.max_entries;
    sh.level_out = p.level_max_out;
    return sh;
}

}  // namespace

void msm_batch_plan(int curve, const MsmJob* specs, size_t njobs, int chunk, MsmBatchPlan& plan) {
    if (curve != 1 && curve != 2) throw std::invalid_argument("curve");
    if (njobs == 0 || njobs > (1u << 24)) throw std::invalid_argument("job count out of range");
    plan.curve = curve;
    plan.njobs = (uint32_t)njobs;
    plan.chunk = (uint32_t)(chunk > 0 ? chunk : 32);
    plan.jobs.assign(specs, specs + njobs);
    uint64_t total_buckets = 0;
    uint32_t min_b = 1u << 17;
    for (auto& j : plan.jobs) {
        const int c = j.c ? (int)j.c : msm_window_bits(curve, j.n, kMsmBatched);
        if (c < 5 || c > 18) throw std::invalid_argument("window bits out of range [5, 18]");
        j.c = (uint32_t)c;
        total_buckets += (uint64_t)(j.table ? 1 : msm_num_windows(c)) << (c - 1);
        const uint64_t span = j.table ? (uint64_t)msm_table_rows(c) * j.table : (uint64_t)j.n;
        if ((j.table && j.pt_base + (uint64_t)j.n > j.table) ||
            (uint64_t)j.tab_base + j.pt_base + span > 0x80000000ull)
            throw std::invalid_argument("job beyond its table or 2^31 points");
        if ((1u << (c - 1)) < min_b) min_b = 1u << (c - 1);
    }
    plan.seg = reduce_segment(total_buckets, min_b);
    std::vector<uint32_t> lv_base, lv_n;  // per window: first segment, segment count
    uint64_t nb = 0, nwin = 0, nseg = 0, blocks = 0;
    size_t entries = 0, max_n = 0;
    for (auto& j : plan.jobs) {
        const int c = (int)j.c;
        const uint32_t Wd = (uint32_t)msm_num_windows(c);                            // digit windows
        const uint32_t W = j.table ? 1u : Wd, B = 1u << (c - 1), t1 = B / plan.seg;  // bucket windows
        j.nwin = W;
        j.win_base = (uint32_t)nwin;
        j.gb_base = (uint32_t)nb;
        j.blk_base = (uint32_t)blocks;
        for (uint32_t w = 0; w < W; ++w) {
            lv_base.push_back((uint32_t)(nseg + (uint64_t)w * t1));
            lv_n.push_back(t1);
        }
        nb += (uint64_t)W * B;
        nwin += W;
        nseg += (uint64_t)W * t1;
        blocks += (j.n + 255) / 256;
        entries += (size_t)j.n * Wd;
        // a bucket holds at most one digit per point and digit window of its job
        const size_t per_bucket = (size_t)j.n * (j.table ? Wd : 1u);
        if (per_bucket > max_n) max_n = per_bucket;

    }
    if (nb >= 0xFFFFFFFFull || blocks >= 0x80000000ull) throw std::invalid_argument("batch too large");
    plan.nb = (uint32_t)nb;
    plan.nwin = (uint32_t)nwin;
    plan.nseg = (uint32_t)nseg;
    plan.dig_blocks = (uint32_t)blocks;
    plan.max_entries = entries;
    // LDS-histogram sort when every job's histogram fits the LDS: parts of
    // >= 8192 points (at most 64 per window), one workgroup per (job, digit
    // window, part)
    {
        std::vector<SortJob> sj(plan.jobs.size());
        uint64_t dig = 0, cnt = 0, items = 0;
        uint32_t max_b = 0;
        bool fits = true;
        for (size_t k = 0; k < plan.jobs.size(); ++k) {
            const MsmJob& j = plan.jobs[k];
            const uint32_t wd = (uint32_t)msm_num_windows((int)j.c), B = 1u << (j.c - 1);
            uint32_t parts = 1;
            while (parts < 64 && (uint64_t)parts * 2 * 8192 <= j.n) parts *= 2;
            SortJob& so = sj[k];
            so.dig_base = dig;
            so.cnt_base = cnt;
            so.item_base = (uint32_t)items;
            so.parts = parts;
            so.per_part = (j.n + parts - 1) / parts;
            so.wd = wd;
            dig += (uint64_t)wd * j.n;
            cnt += (uint64_t)wd * parts * B;
            items += (uint64_t)wd * parts;
            if (B > max_b) max_b = B;
            fits = fits && j.c <= 16;
        }
        plan.lds_sort = fits && items < 0x80000000ull && batch_lds_sort_prepare((size_t)max_b * 4);
        if (plan.lds_sort) {
            plan.sort_items = (uint32_t)items;
            plan.max_b = max_b;
            plan.sort_words = cnt;
            plan.sort_digits = dig;
            plan.d_sjobs.ensure(sj.size() * sizeof(SortJob));
            KZ_HIP(hipMemcpy(plan.d_sjobs.ptr, sj.data(), sj.size() * sizeof(SortJob), hipMemcpyHostToDevice));
        }
    }
    plan.levels = fold_levels(max_n, plan.chunk);
    // reduction levels until every window is down to one part; the last level
    // writes window w's sum at index w
    std::vector<RedLevel> tab;
    plan.level_off.clear();
    plan.level_grid.clear();
    plan.level_max_out = 0;
    for (;;) {
        bool final = true;
        for (uint32_t n : lv_n) final = final && parts_of(n) == 1;
        plan.level_off.push_back((uint32_t)tab.size());
        uint32_t wg = 0;
        for (size_t w = 0; w < lv_n.size(); ++w) {
            const uint32_t parts = parts_of(lv_n[w]);
            RedLevel d;
            d.in_base = lv_base[w];
            d.n_in = lv_n[w];
            d.out_base = final ? (uint32_t)w : wg;
            d.wg_base = wg;
            tab.push_back(d);
            lv_base[w] = d.out_base;
            lv_n[w] = parts;
            wg += parts;
        }
        plan.level_grid.push_back(wg);
        if (final) break;
        if (wg > plan.level_max_out) plan.level_max_out = wg;
    }
    plan.d_jobs.ensure(njobs * sizeof(MsmJob));
    plan.d_levels.ensure(tab.size() * sizeof(RedLevel));
    KZ_HIP(hipMemcpy(plan.d_jobs.ptr, plan.jobs.data(), njobs * sizeof(MsmJob), hipMemcpyHostToDevice));
    KZ_HIP(hipMemcpy(plan.d_levels.ptr, tab.data(), tab.size() * sizeof(RedLevel), hipMemcpyHostToDevice));
}

void msm_batch_reserve(const MsmBatchPlan& plan, MsmWorkspace& ws) {
    reserve(stages_of(plan.curve), batch_shape(plan), ws);
    if (plan.lds_sort) {
        ws.sort_cnt.ensure(plan.sort_words * sizeof(uint32_t));
        ws.sort_pre.ensure(plan.sort_words * sizeof(uint32_t));
        ws.sort_dig.ensure(plan.sort_digits * sizeof(uint16_t));
    }
}

void msm_batch_run(const MsmBatchPlan& plan, const uint32_t* d_pts, const uint8_t* d_inf,
                   const uint32_t* d_scalars, MsmWorkspace& ws, hipStream_t s, uint8_t* d_out) {
    const Stages& st = stages_of(plan.curve);
    const Shape sh = batch_shape(plan);
    WorkspaceUse use(ws, s);
    msm_batch_reserve(plan, ws);
    const MsmJob* d_jobs = plan.d_jobs.as<MsmJob>();
    auto sort = [&] {
        if (plan.lds_sort) {
            batch_lds_sort_dispatch(d_scalars, d_inf, d_jobs, plan.d_sjobs.as<SortJob>(), plan.njobs, plan.dig_blocks,
                                    plan.sort_items, plan.nb, (size_t)plan.max_b * 4, ws.sort_dig.as<uint16_t>(),
                                    ws.sort_cnt.as<uint32_t>(), ws.sort_pre.as<uint32_t>(), ws.counts.as<uint32_t>(),
                                    ws.offsets.as<uint32_t>(), ws.block_sums.as<uint32_t>(), ws.total.as<uint32_t>(),
                                    ws.entries.as<uint32_t>(), s);
            return;
        }
        atomic_sort(sh, [&](bool scatter, uint32_t* cc, uint32_t* entries) {
            digits_multi_dispatch(scatter, d_scalars, d_inf, d_jobs, plan.njobs, plan.dig_blocks, cc, entries, s);
        }, ws, s);
    };
    run_stages(st, sh, sort, &plan, d_jobs, MsmJob{}, plan.njobs, d_pts, ws, s, d_out);
}

// ------------------------------------------------------------ conversions
void points_from_wire(int curve, const uint8_t* d_wire, uint32_t* d_pts, uint8_t* d_inf, size_t n,
                      hipStream_t s) {
    if (n == 0) return;
    if (curve == 3) throw std::invalid_argument("Banderwagon points are read by the validating reader");
    if (curve == 2) launch_points_from_wire_bls(d_wire, d_pts, d_inf, (uint32_t)n, s);
    else if (curve == 1) launch_points_from_wire_bn(d_wire, d_pts, d_inf, (uint32_t)n, s);
    else launch_points_from_wire_secp(d_wire, d_pts, d_inf, (uint32_t)n, s);
    KZ_HIP(hipGetLastError());
}

void compress_on_device(int curve, const uint8_t* d_wire, uint8_t* d_out, hipStream_t s) {
    if (curve == 2) launch_compress_bls(d_wire, d_out, s);
    else if (curve == 1) launch_compress_bn(d_wire, d_out, s);
    else return;  // secp256k1 records leave the compressed field zero
    KZ_HIP(hipGetLastError());
}

}  // namespace kz
