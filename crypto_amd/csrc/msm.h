// Host interface of the HIP Pippenger MSM pipeline.
#pragma once

#include "common.h"

namespace kz {

// Device scratch reused across calls (grow-only).
struct MsmWorkspace {
    DevBuf counts, offsets, cursor, block_sums, total;
    DevBuf entries;
    DevBuf bsum, carry, run_bucket, segA, segB, winres;
    DevBuf out;
};

struct MsmParams {
    int c = 0;       // window bits; 0 = choose from n
    int chunk = 32;  // sorted entries per accumulation thread
};

int msm_default_window_bits(size_t n);
// Number of windows for c-bit signed digits over 256-bit scalars, including
// the extra window that takes the final carry when c divides 256.
inline int msm_num_windows(int c) {
    int w = (256 + c - 1) / c;
    if (w * c == 256) ++w;
    return w;
}

// Bytes of the device output record: affine wire (BE x || y, zero = identity)
// followed by the compressed encoding.
inline size_t msm_out_bytes(int curve) { return curve == 2 ? 96 + 48 : 64 + 32; }
inline size_t point_wire_bytes(int curve) { return curve == 2 ? 96 : 64; }
inline size_t compressed_bytes(int curve) { return curve == 2 ? 48 : 32; }
// Device record sizes in 32-bit words (radix-2^29 limbs, padded to 16 bytes):
// affine points 28 (BLS12-381) / 20 (BN254), XYZZ points 56 / 36.
inline int affine_words(int curve) { return curve == 2 ? 28 : 20; }
inline int xyzz_words(int curve) { return curve == 2 ? 56 : 36; }

// Wire affine points (BE, zero = identity) -> Montgomery affine (2N words each)
// plus per-point identity flags.
void points_from_wire(int curve, const uint8_t* d_wire, uint32_t* d_pts, uint8_t* d_inf, size_t n,
                      hipStream_t s);

// sum s_i P_i.  d_pts: Montgomery affine points; d_inf: identity flags or null;
// d_scalars: n x 32-byte LE.  Writes msm_out_bytes(curve) bytes at d_out.
void msm_device(int curve, const uint32_t* d_pts, const uint8_t* d_inf, const uint32_t* d_scalars, size_t n,
                const MsmParams& prm, MsmWorkspace& ws, hipStream_t s, uint8_t* d_out);

// Opt-in per-stage timing of msm_device (HIP events on the MSM's stream).
enum MsmStage { kStageSort = 0, kStageAccumulate = 1, kStageReduce = 2, kStageCombine = 3, kStageCount = 4 };
void msm_set_profiling(int on);
int msm_collect_profile(double* sums_ms, int* counts, int nstages);

// Compress an affine wire point on the device (one point).
void compress_on_device(int curve, const uint8_t* d_wire, uint8_t* d_out, hipStream_t s);

}  // namespace kz
