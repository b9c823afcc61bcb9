// Host interface of the HIP Pippenger MSM pipeline.
#pragma once

#include <array>
#include <vector>

#include "common.h"

namespace kz {

// Device scratch reused across calls (grow-only).  Successive MSMs over one
// workspace are ordered on the device even when they run on different
// streams: each records `last_use` on its stream, and the next one's stream
// waits for it first (outside graph capture).
struct MsmWorkspace {
    MsmWorkspace() = default;
    MsmWorkspace(const MsmWorkspace&) = delete;
    MsmWorkspace& operator=(const MsmWorkspace&) = delete;
    ~MsmWorkspace() {
        if (last_use) (void)hipEventDestroy(last_use);
    }
    hipEvent_t last_use = nullptr;
    hipStream_t last_stream = nullptr;
    DevBuf counts, offsets, cursor, block_sums, total;
    DevBuf entries;
    DevBuf bsum, carry, run_bucket, segA, segB, winres;
    DevBuf lvA[2], lvB[2];  // reduction levels (ping-pong)
    DevBuf sort_cnt, sort_pre;  // LDS sort: per-(window, part) histograms
    DevBuf glv_pts, glv_inf, glv_sc;  // GLV split: 2n points, flags, 16-byte scalar magnitudes
    DevBuf sort_dig;                  // batch LDS sort: 16-bit digit codes
    DevBuf out;
    DevBuf slices;                    // MSMs beyond one pipeline: slice results and their sum's inputs
    static constexpr int kBuffers = 24;
    // Every buffer becomes a piece of `a`, carved afresh on its next ensure.
    void bind(DevArena* a) {
        for (DevBuf* b : buffers()) b->bind(a);
    }
    // Current device addresses of every buffer: a captured graph that used
    // this workspace is valid only while they are unchanged (a growing
    // buffer is freed and allocated anew).
    void addresses(const void** out) const {
        int k = 0;
        for (const DevBuf* b : const_cast<MsmWorkspace*>(this)->buffers()) out[k++] = b->ptr;
    }

private:
    std::array<DevBuf*, kBuffers> buffers() {
        return {&counts, &offsets, &cursor, &block_sums, &total, &entries, &bsum, &carry, &run_bucket,
                &segA,   &segB,    &winres, &lvA[0],     &lvA[1], &lvB[0],  &lvB[1], &sort_cnt, &sort_pre,
                &glv_pts, &glv_inf, &glv_sc, &sort_dig, &out, &slices};
    }
};

// While an object of this type lives on a thread, MSMs on that thread record
// no profiler events (so they create no HIP events).
struct NoStageTimers {
    NoStageTimers();
    ~NoStageTimers();
};

struct MsmParams {
    int c = 0;         // window bits; 0 = choose from n
    int chunk = 0;     // sorted entries per accumulation thread; 0 = from the entries per bucket
    bool g1 = false;   // caller guarantees every point lies in G1: the GLV split may be used
    int seg = 0;       // first-level reduction segment; 0 = tuned cell (when c = 0) or rule
    // > 0: d_pts is a window table built with window bits c for table_stride
    // points (row w = 2^(c w) P_i at w * table_stride + i, n <= table_stride);
    // every window's digits then share one bucket set, and no window combine
    // (no doubling chain) is left.  d_inf stays the flags of the n points.
    uint32_t table_stride = 0;
};

// Device descriptor of one MSM inside a batch.  A batch runs every job's
// stages together: global bucket ids, windows and outputs are the jobs'
// ranges laid end to end.
struct MsmJob {
    uint64_t sc_base;   // first scalar (index into the batch's scalar array)
    uint32_t n;         // points
    uint32_t pt_base;   // first point (index into the point array)
    uint32_t c;         // window bits
    uint32_t nwin;      // windows
    uint32_t win_base;  // first global window
    uint32_t gb_base;   // first global bucket
    uint32_t blk_base;  // first block of the digit kernels
    uint32_t table;     // > 0: the points are a window table of this row stride
                        // (row w = 2^(c w) P at tab_base + pt_base + w * table
                        // + i); the job's digits then share one window of buckets
    uint32_t tab_base;  // added to every entry's point index (not to the
                        // identity-flag index pt_base + i)
};

// LDS-histogram sort of a batch: per job, the signed digits are written once
// (16-bit codes, window-major: digit w of point i at dig_base + w n + i), then
// one workgroup per (job, digit window, part of the points) counts them in an
// LDS histogram and scatters them through LDS cursors.
struct SortJob {
    uint64_t dig_base;   // first digit code of the job
    uint64_t cnt_base;   // first histogram word: item (w, p) at cnt_base + (w parts + p) B
    uint32_t item_base;  // first workgroup of the job
    uint32_t parts;      // parts per digit window
    uint32_t per_part;   // points per part
    uint32_t wd;         // digit windows
};

// Inputs per workgroup (one part) of a bucket-reduction level.
constexpr uint32_t kReduceFanIn = 128;
// Carry runs of at most this many chunks are folded serially by one thread;
// longer ones go through the pairwise tree.
constexpr uint32_t kFoldShortRun = 32;

// One window's slice of a bucket-reduction level: inputs [in_base, in_base +
// n_in), outputs [out_base, out_base + ceil(n_in / kReduceFanIn)), served by
// the workgroups from wg_base on.
struct RedLevel {
    uint32_t in_base, n_in, out_base, wg_base;
};

int msm_default_window_bits(size_t n);
// Number of windows for c-bit signed digits over `bits`-bit scalars (256;
// 128 for the halves of a GLV split), including the extra window that takes
// the final carry when c divides the bit count.
inline int msm_num_windows(int c, int bits = 256) {
    int w = (bits + c - 1) / c;
    if (w * c == bits) ++w;
    return w;
}

// Bytes of the device output record: affine wire (BE x || y, zero = identity)
// followed by the compressed encoding.
inline size_t msm_out_bytes(int curve) { return curve == 2 ? 96 + 48 : 64 + 32; }
inline size_t point_wire_bytes(int curve) { return curve == 2 ? 96 : 64; }
inline size_t compressed_bytes(int curve) { return curve == 2 ? 48 : 32; }
// Device record sizes in 32-bit words (radix-2^29 limbs, padded to 16 bytes):
// affine points 28 words of data in a 32-word (128-byte) slot for BLS12-381
// and for Banderwagon (x, y, d x y), 20 for BN254 and secp256k1; projective
// points: XYZZ 56 / 36 / 40, Banderwagon's extended (X, Y, T, Z) 36.
inline int affine_words(int curve) { return curve == 2 || curve == 3 ? 32 : 20; }  // record stride
inline int xyzz_words(int curve) { return curve == 2 ? 56 : curve == 1 || curve == 3 ? 36 : 40; }

// Wire affine points (BE, zero = identity) -> Montgomery affine (2N words each)
// plus per-point identity flags.
void points_from_wire(int curve, const uint8_t* d_wire, uint32_t* d_pts, uint8_t* d_inf, size_t n,
                      hipStream_t s);

// sum s_i P_i.  d_pts: Montgomery affine points; d_inf: identity flags or null;
// d_scalars: n x 32-byte LE.  Writes msm_out_bytes(curve) bytes at d_out.
// Beyond kMsmSlice points (not for window tables) it runs slices of that size
// and sums their results.
constexpr size_t kMsmSlice = (size_t)1 << 25;
void msm_device(int curve, const uint32_t* d_pts, const uint8_t* d_inf, const uint32_t* d_scalars, size_t n,
                const MsmParams& prm, MsmWorkspace& ws, hipStream_t s, uint8_t* d_out);

// ------------------------------------------------------------------ batches
// A batch of independent MSMs over one point array and one scalar array, run
// as one pipeline: every stage is launched once for all jobs (about 12
// launches whatever the job count), and the per-job window combines -- each
// a serial chain of ~256 doublings -- run side by side.
struct MsmBatchPlan {
    int curve = 0;
    uint32_t njobs = 0;
    uint32_t nb = 0;          // global buckets
    uint32_t nwin = 0;        // global windows
    uint32_t nseg = 0;        // first-level reduction segments
    uint32_t dig_blocks = 0;  // blocks of the digit kernels
    uint32_t chunk = 32;
    uint32_t seg = 16;        // buckets per first-level reduction segment
    uint32_t levels = 0;      // carry-fold tree levels
    size_t max_entries = 0;
    std::vector<MsmJob> jobs;  // host copy of the descriptors
    // reduction levels: level l's per-window table at d_levels + level_off[l]
    // (nwin entries), launched on level_grid[l] workgroups
    std::vector<uint32_t> level_off, level_grid;
    uint32_t level_max_out = 0;  // largest output count of a non-final level
    // LDS-histogram sort (every job's c <= 16): workgroups, histogram words,
    // digit codes, largest 2^(c-1)
    bool lds_sort = false;
    uint32_t sort_items = 0, max_b = 0;
    size_t sort_words = 0, sort_digits = 0;
    DevBuf d_jobs, d_levels, d_sjobs;
};

// Tuning modes: one MSM, a job of a batch, one MSM over G1 points through
// the GLV split (2n points of 128-bit scalars).
enum MsmMode { kMsmPlain = 0, kMsmBatched = 1, kMsmGlv = 2, kMsmModes = 3 };

// (curve, mode) pairs with a tuning table: every mode of BN254 (1) and
// BLS12-381 (2); secp256k1 (0) runs single MSMs only, plain or GLV.
inline bool msm_tunable(int curve, int mode) {
    return mode >= 0 && mode < kMsmModes && (curve == 1 || curve == 2 || (curve == 0 && mode != kMsmBatched));
}
// A GLV cell holding kMsmGlvOff instead of window bits records that the
// plain MSM beat the split at that size class: G1 MSMs of that class without
// explicit window bits then run the plain path (same result).
constexpr int kMsmGlvOff = 1;
// Window bits used for an n-point MSM in `mode`: the tuning table entry of
// (curve, mode, floor(log2 n)) when set, else msm_default_window_bits(n)
// (of 2n for GLV); for a GLV cell of kMsmGlvOff, the plain MSM's.
int msm_window_bits(int curve, size_t n, int mode);
// False when the GLV cell of n's size class is kMsmGlvOff.
bool msm_glv_enabled(int curve, size_t n);
// Tuned first-level reduction segment of that size class, 0 = built-in rule.
int msm_tuned_segment(int curve, size_t n, int mode);
// Installs window bits: c_by_log[k] for 2^k <= n < 2^(k+1) (0 = default
// rule), k < count <= 32; count = 0 clears them.  Segments are kept.
void msm_set_window_table(int curve, int mode, const int* c_by_log, int count);
// One tuning table cell (lg = floor(log2 n)); c or seg 0 = built-in rule; c
// may be kMsmGlvOff in mode kMsmGlv.
void msm_set_tuned(int curve, int mode, int lg, int c, int seg);
void msm_get_tuned(int curve, int mode, int lg, int& c, int& seg);

// Builds a plan (synchronous upload of the descriptor tables).  Each spec
// gives n, pt_base, sc_base and c (0 = msm_window_bits(curve, n, kMsmBatched)).
void msm_batch_plan(int curve, const MsmJob* specs, size_t njobs, int chunk, MsmBatchPlan& plan);
// Grows ws for the plan; after this, msm_batch_run allocates nothing and
// issues only kernel launches (so it can be captured in a HIP graph).
void msm_batch_reserve(const MsmBatchPlan& plan, MsmWorkspace& ws);
// Job j's output record (msm_out_bytes) goes to d_out + j * msm_out_bytes.
void msm_batch_run(const MsmBatchPlan& plan, const uint32_t* d_pts, const uint8_t* d_inf,
                   const uint32_t* d_scalars, MsmWorkspace& ws, hipStream_t s, uint8_t* d_out);

// Opt-in per-stage timing of msm_device (HIP events on the MSM's stream).
enum MsmStage { kStageSort = 0, kStageAccumulate = 1, kStageReduce = 2, kStageCombine = 3, kStageCount = 4 };
void msm_set_profiling(int on);
// Buckets per first-level reduction segment (power of two <= 16; 0 = the
// built-in rule) for MSMs and batch plans created afterwards.
void msm_set_reduce_segment(int seg);
int msm_collect_profile(double* sums_ms, int* counts, int nstages);

// Window tables (MsmParams::table_stride): rows = windows of c-bit signed
// digits over 256-bit scalars; the table holds rows * n affine records.
inline int msm_table_rows(int c) { return msm_num_windows(c, 256); }
// Window bits for a table of n points: the widest c <= 16 that still gives
// about 32 digits per bucket (rows * n / 2^(c-1) >= 32), at least 5.
int msm_table_window_bits(size_t n);
// Builds the table of n prepared points (d_table: msm_table_rows(c) * n *
// affine_words(curve) words).  Asynchronous on s.
void points_table_build(int curve, const uint32_t* d_pts, const uint8_t* d_inf, size_t n, int c, uint32_t* d_table,
                        hipStream_t s);

// Compress an affine wire point on the device (one point).
void compress_on_device(int curve, const uint8_t* d_wire, uint8_t* d_out, hipStream_t s);

}  // namespace kz
