/*
 * Multi-curve Pippenger MSM, HIP backend for AMD MI355X (gfx950).
 *
 *   R = sum_{i=0..n-1} s_i * P_i
 *
 * Wire ABI, identical to the reference (gpukit/include/kinet/gpukit/multi_pippenger.h:27-39,
 * as the reference code implements it in gpukit/cpp/cpu_reference/multi_pippenger.cpp:267-290
 * and gpukit/curve_traits/bls12_381_g1_first_party_traits.h:58-85):
 *   scalars : n x 32 bytes, little-endian; any 256-bit value (not reduced mod r)
 *   points  : BN254 G1 n x 64 bytes, BLS12-381 G1 n x 96 bytes; affine x || y,
 *             each coordinate big-endian; all-zero bytes = identity;
 *             coordinates are read mod p
 *   result  : same layout as one point; identity = all-zero bytes
 *
 * Results equal sum(s_i * P_i) exactly for points on the curve, so they match
 * the reference byte for byte for every scalar below 2^255 (every s < r).  For
 * scalars >= 2^255 this backend keeps the signed-digit carry that the
 * reference drops out of its last window (multi_pippenger.cpp:118-126).
 *
 * Banderwagon (multi_pippenger.cpp:314-350): points are the 64-byte
 * uncompressed form (affine x || y on the Bandersnatch curve, big-endian) and
 * are checked as Element::deserialize_uncompressed does (banderwagon/cpp/
 * element.cpp:330-355: x, y below q, on the curve, 1 - a x^2 a square);
 * all-zero bytes are the identity; scalars must be canonical, below the group
 * order (fr.cpp:305-328).  Any rejected point or scalar fails the whole call
 * with GPUKIT_ERR_BAD_SIZE, as in the reference.  The result is the exact sum
 * of the given representatives (two-torsion component included) written as
 * affine x || y; the identity, and n == 0, give (0, 1).
 *
 * Return codes are gpukit_status (gpukit.h).
 */
#pragma once
#include <stddef.h>
#include <stdint.h>

#include "kinet/gpukit/gpukit.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef enum gpukit_curve {
    GPUKIT_CURVE_SECP256K1 = 0,   /* 64-byte BE x || y points, like BN254 */
    GPUKIT_CURVE_BN254_G1 = 1,
    GPUKIT_CURVE_BLS12_381_G1 = 2,
    GPUKIT_CURVE_BANDERWAGON = 3, /* 64-byte BE x || y points, validated (see below) */
} gpukit_curve;

/* The HIP backend, a sibling of gpukit_multi_pippenger_{cpu,cuda,metal,wgsl}
 * (gpukit/include/kinet/gpukit/multi_pippenger.h:69-92).  The multi-backend
 * entry gpukit_multi_pippenger stays the reference's dispatcher
 * (gpukit/c-abi/c_multi_pippenger.cpp:9-42): it calls this function for
 * GPUKIT_BACKEND=hip and keeps its CPU fallback on GPUKIT_ERR_NOTIMPL and
 * GPUKIT_ERR_BACKEND.  This library defines no gpukit_multi_pippenger, so it
 * links beside the reference's gpukit without a duplicate symbol.
 *
 * BN254 and secp256k1 points that all lie on the curve (n >= 4096) take the
 * GLV split (cofactor 1: same bytes, faster).  As in the reference, points off
 * the curve are not rejected (BN254, BLS12-381, secp256k1), and their sum is
 * undefined: the result for such input may differ between calls.
 * n == 0 gives the identity.  Above 2^25 points the MSM runs as slices of
 * 2^25 points (sorted bucket entries carry 31-bit indices) whose results are
 * summed on the device, with the same bytes as one pass; n >= 2^32 gives
 * GPUKIT_ERR_BAD_SIZE (32-bit point indices; 2^32 BLS12-381 points would be
 * 412 GB of wire input).  Device failures, out of device memory included,
 * give GPUKIT_ERR_BACKEND (message via kinet_hip_last_error()). */
int gpukit_multi_pippenger_hip(uint32_t curve, const uint8_t* scalars, const uint8_t* points, size_t n,
                               uint8_t* result);

#ifdef __cplusplus
} /* extern "C" */
#endif
