// GLV endomorphism split and G1 membership check (glv.hip).
#pragma once

#include "common.h"

namespace kz {

// BLS12-381: d_pts2 = [P_i, phi(P_i)] (2n prepared records), d_inf2 likewise,
// d_scalars2 = [s1_i, s2_i] with s_i = s1_i + lambda s2_i mod r (128-bit
// halves in 32-byte records).  Valid only for points of G1.
void glv_expand(int curve, const uint32_t* d_pts, const uint8_t* d_inf, const uint32_t* d_scalars, size_t n,
                uint32_t* d_pts2, uint8_t* d_inf2, uint32_t* d_scalars2, hipStream_t s);

// *d_bad |= 1 if a non-identity prepared point is off the curve or, for
// BLS12-381, outside G1 (*d_bad zeroed by the caller).
void check_g1(int curve, const uint32_t* d_pts, const uint8_t* d_inf, size_t n, int* d_bad, hipStream_t s);

}  // namespace kz
