/*
 * Runtime controls and device-pointer entry points of the HIP library.
 *
 * No counterpart in the reference: the reference's GPU drivers take host
 * buffers only (e.g. kzg/gpu/cuda/kzg_driver_cuda.cpp:44-58).  These entries
 * serve callers whose inputs already live in device memory (benchmarks,
 * batching and serving layers).  Device pointers are HIP device addresses;
 * `stream` is a hipStream_t (NULL = the null stream).  Calls are asynchronous
 * on that stream; concurrent calls that share the library's MSM scratch are
 * serialised on the host.
 */
#pragma once
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* 1 if a HIP device is usable, else 0 (reason via kinet_hip_last_error). */
int kinet_hip_available(void);

/* Message of the last failure inside the library (thread-shared, best effort). */
const char* kinet_hip_last_error(void);

/* Wire points (BE affine, zero = identity) -> device form: Montgomery affine
 * records of 28 (BLS12-381) or 20 (BN254) 32-bit words per point (radix-2^29
 * limbs, 16-byte aligned), and a byte per point flagging the identity. */
int kinet_hip_points_prepare(uint32_t curve, const void* d_wire, size_t n, void* d_pts, void* d_inf,
                             void* stream);

/* MSM over prepared points.  d_scalars: n x 32 bytes LE.  window_bits: 0 picks
 * from n, else 5..18.  chunk: sorted bucket entries per accumulation thread,
 * 0 for the default.  Writes the affine wire result (64/96 bytes) followed by
 * its compressed form (32/48 bytes) at d_out. */
int kinet_hip_msm_prepared(uint32_t curve, const void* d_scalars, const void* d_pts, const void* d_inf,
                           size_t n, int window_bits, int chunk, void* d_out, void* stream);

/* Fixed-base batch scalar multiplication: out[i] = k_i * G for n scalars
 * (32-byte LE, any 256-bit value) on the device, written in the point wire
 * format (BE affine x || y, identity = zero bytes).  The table of multiples of
 * G is built on first use per curve and stays resident.  Synthetic MSM inputs
 * and SRS points [tau^i] G come from here.  gpukit_status codes. */
int kinet_hip_g1_fixed_base_mul(uint32_t curve, const void* d_scalars, size_t n, void* d_out_wire, void* stream);

/* Opt-in stage profiler of the MSM pipeline: while enabled, each MSM records
 * HIP events on its own stream around its stages (0 digit sort, 1 bucket
 * accumulation, 2 bucket reduction, 3 window combine + output).  Collect sums
 * the elapsed milliseconds and call counts per stage (waiting for the events)
 * and forgets them. */
void kinet_hip_profile_enable(int on);
int kinet_hip_profile_collect(double* stage_ms, int* stage_calls, int nstages);

/* Measured VALU throughput of the current device, in instructions per second
 * summed over all lanes: op 0 v_mad_u64_u32 (the roofline denominator of the
 * MSM and NTT kernels), 1 v_mul_lo_u32 + add, 2 v_mul_hi_u32 + xor,
 * 3 v_mad_u32_u24, 4 v_fma_f64, 5 add3 + shift.  Returns 0 on success. */
int kinet_hip_valu_peak(int op, double* ops_per_second);

#ifdef __cplusplus
} /* extern "C" */
#endif
