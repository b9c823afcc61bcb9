/*
 * Scalar-field NTT on AMD MI355X (gfx950) for the KZG commitment path.
 *
 * The reference's NTT entries (c-abi/crypto.h:368-369, ntt/c-abi/c_ntt.cpp:81-146)
 * take uint64_t words and cannot carry a 255-bit Fr element; they stay where
 * they are.  These entries add the Fr transform the commit path needs.
 *
 * Field: 0 = BLS12-381 Fr (r = 0x73eda753...00000001), 1 = BN254 Fr.
 * Elements: n x 32 bytes, little-endian canonical (< r), natural order in and
 * out.  n a power of two, 1 <= n <= 2^26 (and <= 2^28 for BN254 by its
 * 2-adicity).
 * Forward:  X_k = sum_j x_j w^(jk),  w = g^((r-1)/n), g = 7 (BLS12-381), 5 (BN254).
 * Inverse:  x_j = n^-1 sum_k X_k w^(-jk).
 * Returns CRYPTO_OK (0); CRYPTO_ERR_INPUT (-1) for a bad field, null pointer,
 * bad n, or an element >= r (elements then untouched); CRYPTO_ERR_BACKEND (-4)
 * if the device fails (message via kinet_hip_last_error()).
 */
#pragma once
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KINET_NTT_FIELD_BLS12_381_FR 0
#define KINET_NTT_FIELD_BN254_FR 1

int ntt_fr_forward(uint32_t field, uint8_t* elems, size_t n);
int ntt_fr_inverse(uint32_t field, uint8_t* elems, size_t n);

/* Device-pointer variant, asynchronous on `stream` (hipStream_t, NULL = null
 * stream).  d_in and d_out may alias; d_tmp is n x 32 bytes of caller scratch;
 * *d_err (device int, zeroed by the caller) becomes nonzero if an input is >= r.
 * Twiddle tables are built on first use per (field, n, direction) and stay
 * resident. */
int kinet_hip_ntt_fr(uint32_t field, int inverse, const void* d_in, void* d_out, void* d_tmp, size_t n, int* d_err,
                     void* stream);

/* `count` transforms of the same size n laid back to back (d_in, d_out, d_tmp:
 * count x n x 32 bytes), one launch per pass for all of them; count <= 65535.
 * Same element format, domain and error flag as kinet_hip_ntt_fr. */
int kinet_hip_ntt_fr_batch(uint32_t field, int inverse, const void* d_in, void* d_out, void* d_tmp, size_t n,
                           size_t count, int* d_err, void* stream);

#ifdef __cplusplus
} /* extern "C" */
#endif
