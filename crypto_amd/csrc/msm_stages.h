// Launchers of the curve-templated MSM kernels.  Each stage lives in its own
// translation unit per curve (msm_<stage>_<curve>.hip) so the builds run in
// parallel; the host pipeline (msm_host.hip) calls them through these.
#pragma once

#include "msm.h"

namespace kz {

#define KZ_DECLARE_STAGES(sfx)                                                                              \
    void launch_points_from_wire_##sfx(const uint8_t* wire, uint32_t* pts, uint8_t* inf, uint32_t n,        \
                                       hipStream_t s);                                                      \
    void launch_compress_##sfx(const uint8_t* wire, uint8_t* out, hipStream_t s);                           \
    void launch_accumulate_##sfx(unsigned grid, hipStream_t s, const uint32_t* entries,                     \
                                 const uint32_t* offsets, uint32_t nb, uint32_t chunk, const uint32_t* pts, \
                                 uint32_t* bsum, uint32_t* carry);                                          \
    void launch_fold_##sfx(hipStream_t s, const uint32_t* offsets, uint32_t nb, uint32_t chunk,             \
                           uint32_t nchunks, uint32_t levels, uint32_t* list, uint32_t* runs, uint32_t* ctl,   \
                           uint32_t* carry);                                                                \
    void launch_reduce_seg_##sfx(unsigned grid, hipStream_t s, const uint32_t* offsets, uint32_t chunk,     \
                                 const uint32_t* bsum, const uint32_t* carry, uint32_t nseg, uint32_t seg,  \
                                 uint32_t* segA, uint32_t* segB);                                           \
    void launch_reduce_level_##sfx(unsigned grid, hipStream_t s, const uint32_t* inA, const uint32_t* inB,  \
                                   const RedLevel* tab, uint32_t ntab, uint32_t uni_n, uint32_t uni_parts,  \
                                   uint32_t lm, uint32_t* outA, uint32_t* outB);                            \
    void launch_combine_##sfx(hipStream_t s, const uint32_t* winres, const MsmJob* jobs, const MsmJob& one,  \
                              uint32_t njobs, uint8_t* out);                                                \
    void launch_table_rows_##sfx(const uint32_t* pts, const uint8_t* inf, uint32_t n, uint32_t c,           \
                                 uint32_t nwin, uint32_t* table, hipStream_t s);

KZ_DECLARE_STAGES(bls)
KZ_DECLARE_STAGES(bn)
KZ_DECLARE_STAGES(secp)
#undef KZ_DECLARE_STAGES

// Banderwagon (curve 3, msm_bw.hip): the MSM stages, and a wire reader that
// validates (bit 0 of *err: a rejected point; bit 1: a scalar not below the
// group order, from bw_check_scalars).  No window tables.
void launch_accumulate_bw(unsigned grid, hipStream_t s, const uint32_t* entries, const uint32_t* offsets, uint32_t nb,
                          uint32_t chunk, const uint32_t* pts, uint32_t* bsum, uint32_t* carry);
void launch_fold_bw(hipStream_t s, const uint32_t* offsets, uint32_t nb, uint32_t chunk, uint32_t nchunks,
                    uint32_t levels, uint32_t* list, uint32_t* runs, uint32_t* ctl, uint32_t* carry);
void launch_reduce_seg_bw(unsigned grid, hipStream_t s, const uint32_t* offsets, uint32_t chunk, const uint32_t* bsum,
                          const uint32_t* carry, uint32_t nseg, uint32_t seg, uint32_t* segA, uint32_t* segB);
void launch_reduce_level_bw(unsigned grid, hipStream_t s, const uint32_t* inA, const uint32_t* inB,
                            const RedLevel* tab, uint32_t ntab, uint32_t uni_n, uint32_t uni_parts, uint32_t lm,
                            uint32_t* outA, uint32_t* outB);
void launch_combine_bw(hipStream_t s, const uint32_t* winres, const MsmJob* jobs, const MsmJob& one, uint32_t njobs,
                       uint8_t* out);
void bw_points_from_wire(const uint8_t* wire, uint32_t* pts, uint8_t* inf, uint32_t n, uint32_t* err, hipStream_t s);
void bw_check_scalars(const uint32_t* scalars, uint32_t n, uint32_t* err, hipStream_t s);
void bw_identity_record(uint8_t* out, hipStream_t s);

}  // namespace kz
