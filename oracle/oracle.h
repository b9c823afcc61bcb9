/*
 * CPU oracle for the KZG-commitment hot path (G1 MSM over BLS12-381 / BN254
 * and the scalar-field NTT).
 *
 * THIS IS TEST INFRASTRUCTURE, NOT PRODUCT CODE.  Only the tests, the smoke
 * check and the bench's CPU-baseline leg load it, and only as the checker.
 * The product path (crypto_amd/, the HIP library) never links or calls it.
 *
 * It restates, in plain C99 over 64-bit limbs, the reference's CPU algorithms:
 *   - Montgomery CIOS field arithmetic   bls/cpp/bls12_381_fp.hpp:149-189,
 *                                        bn254/cpp/bn254_fp.hpp (same CIOS)
 *   - Jacobian G1 dbl-2009-l / add-2007-bl
 *                                        bls/cpp/bls12_381_g1.hpp:74-151,
 *                                        bn254/cpp/bn254_g1.hpp:66-170
 *   - signed-digit Pippenger + best_c    gpukit/cpp/cpu_reference/multi_pippenger.cpp:67-181
 *   - wire adapters (BE affine, all-zero = identity)
 *                                        gpukit/curve_traits/bls12_381_g1_first_party_traits.h:58-85,
 *                                        gpukit/cpp/cpu_reference/multi_pippenger.cpp:267-290
 *   - radix-2 NTT loop structure         ntt/cpp/ntt.cpp:97-172 (restated over Fr)
 *
 * Deliberate divergence (documented): the reference drops the signed-digit
 * carry out of the last window (multi_pippenger.cpp:118-126,161-165), which is
 * wrong for some scalars >= 2^255.  The oracle keeps that carry in an extra
 * window, so it equals sum(s_i * P_i) for every 256-bit scalar; for s < 2^255
 * (every s < r) both agree byte for byte.
 *
 * Wire conventions (same as the reference ABI, multi_pippenger.h:27-39):
 *   scalars  32-byte little-endian; points BE affine x||y (BN254 64 B,
 *   BLS12-381 96 B); identity = all-zero bytes.
 *   Fr elements for the NTT: 32-byte little-endian canonical.
 */
#ifndef KINET_ORACLE_H
#define KINET_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Curve ids follow gpukit_curve (multi_pippenger.h:50-55). */
#define ORC_CURVE_BN254     1
#define ORC_CURVE_BLS12_381 2

/* Field ids. */
#define ORC_FIELD_BLS_FP 0
#define ORC_FIELD_BLS_FR 1
#define ORC_FIELD_BN_FP  2
#define ORC_FIELD_BN_FR  3
#define ORC_FIELD_BW_FR  4   /* Banderwagon scalar field (banderwagon/cpp/fr.cpp:11-16) */

/* ---- field probes (canonical little-endian bytes in and out) ---- */
int  orc_field_nbytes(int field);
void orc_field_modulus(int field, uint8_t* out_le);
/* op: 0 add, 1 sub, 2 mul, 3 inverse (b ignored), 4 sqrt (returns -1 if none) */
int  orc_field_op(int field, int op, const uint8_t* a_le, const uint8_t* b_le, uint8_t* out_le);
/* Montgomery constants as derived from p (R mod p, R^2 mod p, -p^-1 mod 2^64). */
void orc_field_mont_consts(int field, uint8_t* r1_le, uint8_t* r2_le, uint64_t* pinv);

/* ---- G1 ---- */
int  orc_point_bytes(int curve);               /* 64 or 96 */
int  orc_compressed_bytes(int curve);          /* 32 or 48 */
void orc_g1_generator(int curve, uint8_t* out_wire);
/* out = k * G, k 32-byte LE (any 256-bit value). */
void orc_g1_mul_gen(int curve, const uint8_t* k_le, uint8_t* out_wire);
/* out = k * P (wire in/out). */
void orc_g1_mul(int curve, const uint8_t* p_wire, const uint8_t* k_le, uint8_t* out_wire);
void orc_g1_add(int curve, const uint8_t* a_wire, const uint8_t* b_wire, uint8_t* out_wire);
void orc_g1_dbl(int curve, const uint8_t* a_wire, uint8_t* out_wire);
int  orc_g1_on_curve(int curve, const uint8_t* p_wire);
/* Compression.  BLS12-381: ZCash/IETF 48-byte format.  BN254: 32-byte BE x with
 * the two top bits 0b10 (smaller y), 0b11 (larger y), 0b01 (infinity). */
void orc_g1_compress(int curve, const uint8_t* p_wire, uint8_t* out);
int  orc_g1_decompress(int curve, const uint8_t* in, uint8_t* out_wire);

/* ---- MSM ---- */
int  orc_best_c(size_t n);                     /* multi_pippenger.cpp:76-88 */
/* sum(s_i * P_i) via the signed-digit Pippenger of the reference. */
int  orc_msm(int curve, const uint8_t* scalars, const uint8_t* points, size_t n, uint8_t* out_wire);
/* Same sum by the independent method sum(s_i * P_i) with double-and-add per term. */
int  orc_msm_naive(int curve, const uint8_t* scalars, const uint8_t* points, size_t n, uint8_t* out_wire);
/* Multi-threaded Pippenger: windows split over `threads` POSIX threads. */
int  orc_msm_mt(int curve, const uint8_t* scalars, const uint8_t* points, size_t n, uint8_t* out_wire, int threads);

/* ---- Fr NTT (natural order in and out, X_k = sum_j x_j w^{jk}) ---- */
/* field: 0 = BLS12-381 Fr, 1 = BN254 Fr.  Returns -1 on bad n or non-canonical input. */
int  orc_ntt(int fr, uint8_t* elems, size_t n, int inverse);
int  orc_dft_naive(int fr, const uint8_t* in, uint8_t* out, size_t n, int inverse);
/* Primitive n-th root of unity, canonical LE. */
int  orc_root_of_unity(int fr, size_t n, uint8_t* out_le);

/* ---- KZG ---- */
/* commit = compress(MSM(srs, iNTT(evals))). */
int  orc_kzg_commit(int curve, const uint8_t* srs_wire, const uint8_t* evals, size_t n, uint8_t* out_compressed);
/* closed form: compress([f(tau)] G) with f = iNTT(evals), Horner at tau. */
int  orc_kzg_commit_closed_form(int curve, const uint8_t* tau_le, const uint8_t* evals, size_t n, uint8_t* out_compressed);

/* ---- O(n) closed forms over Fr (fr: 0 = BLS12-381, 1 = BN254; 2 = Banderwagon, dot only) ---- */
/* sum a_i b_i mod r (a_i, b_i any 256-bit LE values). */
int  orc_fr_dot(int fr, const uint8_t* a, const uint8_t* b, size_t n, uint8_t* out_le);
/* sum c_i x^i (canonical inputs; -1 otherwise). */
int  orc_fr_eval(int fr, const uint8_t* coeffs, size_t n, const uint8_t* x_le, uint8_t* out_le);
/* f(tau) from f(w^j), j < n, by the barycentric formula; -1 if tau is in the
 * domain, n is not a supported power of two or an input is not canonical. */
int  orc_fr_lagrange(int fr, const uint8_t* evals, size_t n, const uint8_t* tau_le, uint8_t* out_le);
/* Synthetic division f = (X - z) q + y (kzg/gpu/metal/kzg.metal:201-251):
 * q_out n - 1 elements, y_out = f(z); -1 for n = 0 or non-canonical input. */
int  orc_fr_divide_linear(int fr, const uint8_t* coeffs, size_t n, const uint8_t* z_le, uint8_t* q_out,
                          uint8_t* y_out);

/* ---- Banderwagon (curve 3 of the multi-curve entry) ----
 * Twisted Edwards a x^2 + y^2 = 1 + d x^2 y^2 over BLS12-381 Fr, a = -5
 * (banderwagon/cpp/element.hpp:6-22), projective (X : Y : Z) with the
 * reference's add-2008-bbjlp / dbl-2008-bbjlp (element.cpp:145-209).  Points on
 * the wire: 64 bytes, affine x || y big-endian (serialize_uncompressed,
 * element.cpp:286-300); all-zero bytes = identity on input
 * (multi_pippenger.cpp:324-331); the identity is written as (0, 1).
 * Compressed: 32-byte BE x, negated when y is not lexicographically largest
 * (element.cpp:266-284).  Scalars: 32-byte LE, canonical below the scalar field
 * order (fr.cpp:305-328). */
#define ORC_CURVE_BANDERWAGON 3
void orc_bw_generator(uint8_t* out64);
/* 1 if the 64 wire bytes pass deserialize_uncompressed (element.cpp:330-355). */
int  orc_bw_check(const uint8_t* p64);
/* out = k P for any 256-bit LE k (double-and-add; the exact integer multiple). */
void orc_bw_mul(const uint8_t* p64, const uint8_t* k_le, uint8_t* out64);
void orc_bw_add(const uint8_t* a64, const uint8_t* b64, uint8_t* out64);
void orc_bw_dbl(const uint8_t* a64, uint8_t* out64);
void orc_bw_compress(const uint8_t* p64, uint8_t* out32);
/* The reference's Banderwagon MSM (multi_pippenger.cpp:314-350 into
 * multiexp.cpp:36-157): 0, or -2 (GPUKIT_ERR_BAD_SIZE) for a point or scalar
 * the reference rejects. */
int  orc_bw_msm(const uint8_t* scalars, const uint8_t* points, size_t n, uint8_t* out64);
/* Same sum by double-and-add per term (no Pippenger). */
int  orc_bw_msm_naive(const uint8_t* scalars, const uint8_t* points, size_t n, uint8_t* out64);
/* points k_i G (k_i any 256-bit LE). */
void orc_bw_gen_points(const uint8_t* ks, size_t n, uint8_t* out_wire);

/* ---- seeded inputs ---- */
uint64_t orc_splitmix64(uint64_t* state);
/* n scalars uniform in [0, r) by rejection over masked 4x64-bit draws. */
void orc_gen_scalars(int curve, uint64_t seed, size_t n, uint8_t* out);
/* tau^i for i < n (canonical LE Fr). */
void orc_powers(int curve, const uint8_t* tau_le, size_t n, uint8_t* out);
/* points k_i * G (slow: one scalar multiplication each). */
void orc_gen_points(int curve, const uint8_t* ks, size_t n, uint8_t* out_wire);
/* points (k0 + i*d) * G for i < n (fast: one addition each, batch-normalised). */
void orc_gen_points_progression(int curve, const uint8_t* k0_le, const uint8_t* d_le, size_t n, uint8_t* out_wire);

#ifdef __cplusplus
}
#endif
#endif
