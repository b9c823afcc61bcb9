/*
 * Montgomery batch inversion, HIP backend for AMD MI355X (gfx950).
 *
 *   out[i] = in[i]^-1 mod p
 *
 * Siblings of gpukit_batch_inv_<field>_fp_{cpu,metal,cuda,wgsl}
 * (gpukit/include/kinet/gpukit/batch_inversion.h:25-40), with the contract
 * that header documents (batch_inversion.h:1-15): little-endian elements of
 * 32 bytes (secp256k1, BN254) or 48 bytes (BLS12-381), caller-owned host
 * buffers of identical layout; GPUKIT_OK, or GPUKIT_ERR_BAD_SIZE when an
 * element is zero (out is then left untouched).  Elements are read as their
 * residues mod p (any value that fits the record), so a nonzero multiple of
 * p is a zero element; outputs are canonical (< p).  in and out may be the
 * same buffer.  n = 0 returns GPUKIT_OK.  The reference's CPU form does not
 * return inverses (its generic reduction, batch_inversion.cpp:110-141 and
 * 263-289, subtracts p 2^k without a comparison), so results follow the
 * header's arithmetic rather than that form's bytes.
 * Device-resident data: kinet_hip_batch_inv_fp (kinet/crypto/hip_runtime.h).
 */
#pragma once
#include <stddef.h>
#include <stdint.h>

#include "kinet/gpukit/gpukit.h"

#ifdef __cplusplus
extern "C" {
#endif

/* replaces gpukit_batch_inv_secp256k1_fp_cuda (batch_inversion.h:27, a NOTIMPL stub) */
int gpukit_batch_inv_secp256k1_fp_hip(const uint8_t* in, uint8_t* out, size_t n);
/* replaces gpukit_batch_inv_bn254_fp_cuda (batch_inversion.h:33) */
int gpukit_batch_inv_bn254_fp_hip(const uint8_t* in, uint8_t* out, size_t n);
/* replaces gpukit_batch_inv_bls12_381_fp_cuda (batch_inversion.h:39) */
int gpukit_batch_inv_bls12_381_fp_hip(const uint8_t* in, uint8_t* out, size_t n);

#ifdef __cplusplus
}
#endif
