/* EIP-2537 G1 multi-scalar multiplication on the HIP backend.
 *
 * Replaces the reference's
 *   int bls12_381_g1_msm(uint8_t rx[64], uint8_t ry[64], const uint8_t* xycs, size_t size)
 * (bls/c-abi/c_bls_pairing.h:153-156, body bls/cpp/bls.cpp:194-245) with the
 * same name, arguments and return convention.  Host buffers, synchronous.
 */
#ifndef KINET_CRYPTO_BLS_H
#define KINET_CRYPTO_BLS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* xycs: size / 160 tuples x[64] || y[64] || c[32].  x and y are BLS12-381 Fp
 * elements as 64-byte big-endian integers (so the first 16 bytes are zero and
 * the value is below p); (0, 0) is the point at infinity.  c is a 256-bit
 * big-endian scalar, used in full.  Every point must lie on the curve and in
 * G1.  Writes sum c_k P_k as affine x, y in the same 64-byte encoding ((0, 0)
 * for infinity, also when size is 0).  Returns 1 on success, 0 on any failure
 * (encoding, curve or subgroup check, size not a multiple of 160, no usable
 * device -- there is no CPU route); rx and ry are written only on success. */
int bls12_381_g1_msm(uint8_t rx[64], uint8_t ry[64], const uint8_t* xycs, size_t size);

#ifdef __cplusplus
}
#endif

#endif
