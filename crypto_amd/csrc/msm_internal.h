// Internal host-side declarations shared by the MSM translation units.
#pragma once

#include <stdexcept>

#include "msm.h"

namespace kz {

// Signed-digit count (scatter = false) or scatter (scatter = true) pass.
void digits_dispatch(int c, bool scatter, const uint32_t* d_scalars, const uint8_t* d_inf, uint32_t n,
                     uint32_t* counts_or_cursor, uint32_t* entries, uint32_t* ebucket, hipStream_t s);

// Per-curve pipelines (msm_bls.hip, msm_bn.hip).
void msm_device_bls(const uint32_t* d_pts, const uint8_t* d_inf, const uint32_t* d_scalars, size_t n,
                    const MsmParams& prm, MsmWorkspace& ws, hipStream_t s, uint8_t* d_out);
void msm_device_bn(const uint32_t* d_pts, const uint8_t* d_inf, const uint32_t* d_scalars, size_t n,
                   const MsmParams& prm, MsmWorkspace& ws, hipStream_t s, uint8_t* d_out);
void points_from_wire_bls(const uint8_t* d_wire, uint32_t* d_pts, uint8_t* d_inf, size_t n, hipStream_t s);
void points_from_wire_bn(const uint8_t* d_wire, uint32_t* d_pts, uint8_t* d_inf, size_t n, hipStream_t s);
void compress_bls(const uint8_t* d_wire, uint8_t* d_out, hipStream_t s);
void compress_bn(const uint8_t* d_wire, uint8_t* d_out, hipStream_t s);

}  // namespace kz
