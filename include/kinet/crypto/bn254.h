/*
 * BN254 on the HIP backend: unit operations in the reference's batch layout.
 *
 * HIP counterparts of the reference's BN254 CUDA driver entries
 * kinet_bn254_cuda_{available, g1_add, g1_mul, fp_mul}
 * (bn254/gpu/cuda/bn254_driver_cuda.h:13-28), with the same arguments, layout
 * and return codes.  They run the device field and G1 arithmetic of the MSM
 * one operation per element, so the reference's determinism harness
 * (bn254/test/bn254_gpu_determinism_test.cpp) can drive them like the CUDA
 * and WebGPU drivers.  BN254 MSMs go through gpukit_multi_pippenger_hip
 * (curve 1, kinet/gpukit/multi_pippenger.h) and the KZG entries (kzg.h).
 *
 * Layout: an affine point is 9 x u64 little-endian, x[4] || y[4] || inf
 * (inf != 0: the identity; outputs write x = y = 0, inf = 1 for it); every
 * coordinate in Montgomery form x 2^256 mod p (bn254/cpp/bn254_fp.hpp), and
 * canonical (< p).  A scalar is 4 x u64 LE, any 256-bit value.  An Fp element
 * is 4 x u64 LE in Montgomery form.  Points must lie on the curve.
 * Host buffers, synchronous; n elements each.  Returns 0, or -1 when no HIP
 * device is usable or the device fails (message via kinet_hip_last_error()).
 */
#ifndef KINET_CRYPTO_BN254_H
#define KINET_CRYPTO_BN254_H

#ifdef __cplusplus
extern "C" {
#endif

/* 1 if a HIP device is usable, else 0. */
int kinet_bn254_hip_available(void);

/* out[i] = a[i] + b[i] in G1 (doubling and inverse pairs included). */
int kinet_bn254_hip_g1_add(const void* a, const void* b, void* out, unsigned n);

/* out[i] = scalars[i] * points[i] in G1. */
int kinet_bn254_hip_g1_mul(const void* points, const void* scalars, void* out, unsigned n);

/* out[i] = a[i] * b[i] * 2^-256 mod p: the Montgomery product, so Montgomery
 * inputs give the Montgomery form of the product. */
int kinet_bn254_hip_fp_mul(const void* a, const void* b, void* out, unsigned n);

#ifdef __cplusplus
}
#endif

#endif
