/*
 * KZG polynomial commitments on AMD MI355X (gfx950): C = [f(tau)] G with
 * f given by its evaluations, computed as C = MSM(SRS, iNTT(evals)).
 *
 * The reference declares kzg_blob_to_commit (c-abi/crypto.h:238), which
 * delegates to c-kzg-4844 (absent, and its translation unit does not compile:
 * kzg/cpp/kzg_blob.cpp:18).  Its documentation describes, but does not
 * implement, an SRS-handle API (docs/content/docs/api/kzg.mdx:289-331:
 * kzg_srs_t, kzg_commit).  These entries implement that API on the device;
 * names and conventions follow it and c-abi/crypto.h.
 *
 * Curves: 1 = BN254 G1, 2 = BLS12-381 G1 (gpukit_curve numbering).
 * SRS points: monomial basis [tau^i] G, wire format of gpukit_multi_pippenger
 * (BE affine x || y, 64 or 96 bytes).  They stay resident on the device.
 * Evaluations: n x 32 bytes LE canonical Fr (< r), f(w^j) for the n-th root of
 * unity w of ntt.h (natural order), n a power of two <= SRS size.
 * Commitments: compressed G1 -- BLS12-381 48 bytes (ZCash/IETF format: 0x80
 * compressed flag, 0x40 infinity, 0x20 larger y); BN254 32 bytes (big-endian
 * x; top bits 0b10 smaller y, 0b11 larger y, 0b01 infinity).
 * Returns CRYPTO_OK (0) or CRYPTO_ERR_* (c-abi/crypto.h:60-66): INPUT for bad
 * arguments or non-canonical evaluations, BACKEND for device failures.
 */
#pragma once
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct kzg_srs kzg_srs; /* opaque: device-resident SRS points */
typedef struct kzg_ctx kzg_ctx; /* opaque: a stream plus its scratch */

/* SRS from n wire points (host memory). */
int kzg_srs_create(uint32_t curve, const uint8_t* points, size_t n, kzg_srs** out);
/* Synthetic SRS [tau^i] G, i < n, generated on the device (tau: 32-byte LE, < r). */
int kzg_srs_generate(uint32_t curve, const uint8_t tau[32], size_t n, kzg_srs** out);
void kzg_srs_destroy(kzg_srs* srs);
size_t kzg_srs_size(const kzg_srs* srs);
uint32_t kzg_srs_curve(const kzg_srs* srs);
/* Adds a window table of the SRS on the device for window bits c: rows
 * w = 0..R-1 hold 2^(c w) [tau^i] G (R = windows of c-bit digits over 256
 * bits, 17 at c = 16).  An SRS may hold tables for several c (one call per c;
 * a c already present is a no-op; all tables are rebuilt into one buffer).
 * Every later MSM against the SRS uses them: single commits (kzg_commit,
 * kzg_commit_coeffs, kzg_msm, kzg_commit_async) take the table whose c is
 * nearest to the window width the library picks for their size; batched and
 * served commits take, per request, a table within 2 bits of that same
 * width, else one within 2 bits of the built-in width floor(log2 n) - 4,
 * else ordinary windows.  A table MSM reduces one
 * window and combines none.  Costs R * size * 128 (BLS12-381) or 80 (BN254)
 * bytes of device memory per table: 9.1 GB for a 2^22-point BLS12-381 SRS at
 * c = 16 (kzg_srs_table_bytes gives the exact figure afterwards;
 * kinet_hip_table_bytes the size of one table beforehand).  window_bits 0 = the library's choice for the SRS size, else 5..18;
 * all tables together must stay under 2^31 points.  Synchronous; call it
 * before the SRS is shared with other threads.  Results are unchanged byte
 * for byte.  No counterpart in the reference. */
int kzg_srs_precompute(kzg_srs* srs, int window_bits);
/* Device bytes of the SRS's window tables (0 = none). */
size_t kzg_srs_table_bytes(const kzg_srs* srs);
/* Copies points [first, first + count) back to the host in wire format. */
int kzg_srs_export(const kzg_srs* srs, size_t first, size_t count, uint8_t* points_out);

/* Commitment to the polynomial with the given evaluations (host memory, synchronous). */
int kzg_commit(const kzg_srs* srs, const uint8_t* evals, size_t n, uint8_t* commit_out);
/* Opening proof of f at z, f given as for kzg_commit (n evaluations on the
 * NTT domain, n a power of two >= 2): y = f(z) (32 bytes LE canonical) and
 * proof = [q(tau)] G for q(X) = (f(X) - y) / (X - z), compressed like a
 * commitment.  z: 32 bytes LE canonical, outside the domain (a domain point
 * gives CRYPTO_ERR_INPUT).  The reference's kzg_commit_to_proof
 * (c-abi/crypto.h:239) computes the same opening for an EIP-4844 blob over
 * its trusted setup; this one works over a kzg_srs handle. */
int kzg_open(const kzg_srs* srs, const uint8_t* evals, size_t n, const uint8_t z[32], uint8_t* proof_out,
             uint8_t y_out[32]);
/* Commitment from coefficients c_i (f(X) = sum c_i X^i): MSM + compression
 * only, the form documented for the reference's planned C API
 * (docs/content/docs/api/kzg.mdx:289-331, `kzg_commit(srs, coeffs, ...)`). */
int kzg_commit_coeffs(const kzg_srs* srs, const uint8_t* coeffs, size_t n, uint8_t* commit_out);
/* Opening from coefficients, the documented `kzg_open(srs, coeffs, num_coeffs,
 * z, proof_out, y_out)` (docs/content/docs/api/kzg.mdx:305-312): y = f(z)
 * (32 bytes LE canonical) and proof = [q(tau)] G, compressed like a
 * commitment, for the quotient q(X) = (f(X) - y) / (X - z) by synthetic
 * division (the reference's kzg/gpu/metal/kzg.metal:216-251).  Any
 * 1 <= n <= SRS size and any canonical z, domain points included; n = 1 gives
 * the identity as proof.  CRYPTO_ERR_INPUT for bad arguments or non-canonical
 * coefficients or z. */
int kzg_open_coeffs(const kzg_srs* srs, const uint8_t* coeffs, size_t n, const uint8_t z[32], uint8_t* proof_out,
                    uint8_t y_out[32]);
/* Alias of kzg_srs_destroy under the documented name (kzg.mdx: kzg_srs_free). */
void kzg_srs_free(kzg_srs* srs);

/* count independent commitments, each against its own SRS; commits_out holds
 * count compressed points back to back (48 or 32 bytes each; every SRS of one
 * call must be on the same curve). */
int kzg_commit_batch(const kzg_srs* const* srs, const uint8_t* const* evals, const size_t* n, size_t count,
                     uint8_t* commits_out);
/* MSM of n scalars (32-byte LE, host memory) against the first n SRS points,
 * result in wire format (64 or 96 bytes). */
int kzg_msm(const kzg_srs* srs, const uint8_t* scalars, size_t n, uint8_t* result);

/* ---------------------------------------------------------- EIP-4844 blobs
 * The reference's blob entries (c-abi/crypto.h:238-239, c-kzg-4844 v2.1.7
 * behind them, kzg/cpp/kzg_blob.cpp:69-125), same names, arguments and
 * return codes.  A blob is 4096 BLS12-381 Fr elements, 32-byte big-endian,
 * each < r (else CRYPTO_ERR_INPUT); they are the evaluations of f on the
 * domain w^brp(i), w = 7^((r - 1) / 4096), brp the 12-bit bit reversal.
 *   kzg_blob_to_commit:  commit = [f(tau)] G1, 48-byte compressed
 *   kzg_commit_to_proof: y = f(z) (32-byte big-endian) and the proof
 *                        [(f(tau) - y) / (tau - z)] G1, for any canonical z
 *                        (z >= r: CRYPTO_ERR_INPUT), z in the domain included
 * Both need a trusted setup: the reference embeds the mainnet file at
 * configure time; here it is loaded by kzg_load_trusted_setup_file, or on
 * first use from the file named by the environment variable
 * KINET_KZG_TRUSTED_SETUP.  Without one they return CRYPTO_ERR_BACKEND. */
int kzg_blob_to_commit(const uint8_t blob[131072], uint8_t commit[48]);
int kzg_commit_to_proof(const uint8_t blob[131072], const uint8_t z[32], uint8_t proof[48], uint8_t y[32]);

/* Loads a trusted setup in c-kzg-4844's text format (the file the reference
 * embeds, trusted_setup.txt): "4096", the G2 point count, 4096 compressed G1
 * points in Lagrange form (natural order, hex), the compressed G2 points,
 * then -- in files of c-kzg 2.x -- 4096 compressed G1 points in monomial
 * form.  G1 points are decompressed and checked to lie in G1 on the device;
 * G2 points are checked for count and encoding only (verification is not
 * provided).  Replaces a loaded setup.  CRYPTO_ERR_INPUT for a malformed or
 * invalid file (the old setup is kept), CRYPTO_ERR_BACKEND for device
 * failures.  Names after c-kzg's load_trusted_setup_file / free_trusted_setup
 * (called at kzg/cpp/kzg_blob.cpp:36-47). */
int kzg_load_trusted_setup_file(const char* path);
void kzg_free_trusted_setup(void);

/* Contexts for device-resident, asynchronous use (batching, serving): each
 * owns scratch and runs on `stream` (hipStream_t; NULL = null stream).  Calls
 * on one context are ordered; different contexts may run concurrently. */
int kzg_ctx_create(void* stream, kzg_ctx** out);
void kzg_ctx_destroy(kzg_ctx* ctx);
/* d_evals: n x 32 bytes on the device; d_out receives the wire point followed
 * by its compressed form (96 + 48 or 64 + 32 bytes).  *d_err (device int,
 * zeroed by the caller) becomes nonzero if an evaluation is >= r.
 * window_bits: 0 = default. */
int kzg_commit_async(kzg_ctx* ctx, const kzg_srs* srs, const void* d_evals, size_t n, int window_bits,
                     void* d_out, int* d_err);

#ifdef __cplusplus
} /* extern "C" */
#endif
