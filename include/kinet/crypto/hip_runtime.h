/*
 * Runtime controls and device-pointer entry points of the HIP library.
 *
 * No counterpart in the reference: the reference's GPU drivers take host
 * buffers only (e.g. kzg/gpu/cuda/kzg_driver_cuda.cpp:44-58).  These entries
 * serve callers whose inputs already live in device memory (benchmarks,
 * batching and serving layers).  Device pointers are HIP device addresses;
 * `stream` is a hipStream_t (NULL = the null stream).  Calls are asynchronous
 * on that stream.  The MSM entries without a plan or context of their own
 * (kinet_hip_msm_prepared, _prepared_g1, _msm_table) share one scratch
 * workspace: the host enqueues them one at a time, and each call's stream
 * waits (hipStreamWaitEvent) for the previous user of that workspace, so calls
 * on different streams run one after the other on the device and never
 * overwrite each other's scratch.
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
 * records of 32 (BLS12-381: 28 words of radix-2^29 limbs in a 128-byte slot;
 * Banderwagon: x, y and d x y, 27 words in a 128-byte slot) or 20 (BN254,
 * secp256k1) 32-bit words per point, 16-byte aligned, and a byte per point
 * flagging the identity.  d_pts must hold kinet_hip_points_bytes(curve, n)
 * bytes and d_inf n bytes.  Banderwagon (curve 3) points are validated as
 * gpukit_multi_pippenger_hip validates them: this call then synchronises the
 * stream and returns GPUKIT_ERR_BAD_SIZE if any point is rejected. */
int kinet_hip_points_prepare(uint32_t curve, const void* d_wire, size_t n, void* d_pts, void* d_inf,
                             void* stream);

/* Device bytes of n prepared points of `curve` (the d_pts size of
 * kinet_hip_points_prepare; 0 for an unknown curve).  Authoritative: record
 * sizes follow the device layout, which callers should not hard-code. */
size_t kinet_hip_points_bytes(uint32_t curve, size_t n);

/* MSM over prepared points.  d_scalars: n x 32 bytes LE.  window_bits: 0 picks
 * from n, else 5..18.  chunk: sorted bucket entries per accumulation thread,
 * 0 for the default.  Writes the affine wire result (64/96 bytes) followed by
 * its compressed form (32/48 bytes) at d_out.  Curves 0-3; for Banderwagon the
 * scalars are used as 256-bit integers (the host entry checks that they are
 * below the group order), the identity is written as (0, 1), and the
 * compressed form is the reference's serialize_compressed. */
int kinet_hip_msm_prepared(uint32_t curve, const void* d_scalars, const void* d_pts, const void* d_inf,
                           size_t n, int window_bits, int chunk, void* d_out, void* stream);

/* kinet_hip_msm_prepared for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  Same arguments and output
 * record, with the built-in window rule (c = floor(log2 n) - 4 in [5, 16]) and
 * the default chunking; the bytes equal kinet_hip_msm_prepared's.  Its
 * scratch is a static 64 MiB device array that each device holds until the
 * process ends, so n is limited to 2^16 points (GPUKIT_ERR_BAD_SIZE beyond);
 * the call makes no hipMalloc, hipFree or other allocation (the first call on
 * a device looks up the array's address and creates one ordering event).
 * Calls on one device share the array: the host serialises them and each
 * call's stream waits for the previous call, except under graph capture,
 * where overlapping replays and calls on one device race. */
int kinet_hip_msm_prepared_noalloc(uint32_t curve, const void* d_scalars, const void* d_pts, const void* d_inf,
                                   size_t n, void* d_out, void* stream);

/* The same MSM for points the caller guarantees to lie in G1 (for example
 * checked once by kinet_hip_points_check_g1, or generated as multiples of the
 * generator; for BN254, cofactor 1, every curve point).  It then uses the GLV
 * endomorphism phi(x, y) = (beta x, y) = lambda (x, y): each 256-bit scalar is
 * split as s = k1 + lambda k2 mod r with |k1|, |k2| < 2^127 (BLS12-381: s mod
 * r divided by the 128-bit lambda, then moved by r = lambda^2 + lambda + 1
 * and by lambda to signed parts; BN254: a signed split over a short lattice
 * basis) or < 2^128 (secp256k1, curve 0: the same kind of lattice split; its
 * group is the whole curve, so every curve point qualifies), the signs moved
 * onto the points, giving a 2n-point MSM over half the windows.  Results equal
 * kinet_hip_msm_prepared's for points of G1; for other points they are
 * unspecified.  Curves 0, 1, 2. */
int kinet_hip_msm_prepared_g1(uint32_t curve, const void* d_scalars, const void* d_pts, const void* d_inf,
                              size_t n, int window_bits, int chunk, void* d_out, void* stream);

/* ------------------------------------------------------------- window tables
 * Fixed point sets (a KZG SRS) can trade HBM for the doubling chain: a window
 * table of n prepared points for window bits c holds rows w = 0 .. R-1
 * (R = windows of c-bit signed digits over 256-bit scalars: ceil(256 / c),
 * plus one when c divides 256) with row w = 2^(c w) P_i at record w n + i.
 * An MSM over the table sends the digit of every window to one shared set of
 * 2^(c-1) buckets, so it reduces one window instead of R and needs no window
 * combine; the accumulation work is unchanged.  Any curve points qualify (no
 * subgroup condition).  Results equal kinet_hip_msm_prepared's byte for byte.
 *
 * Window bits the library picks for a table of n points (the widest c <= 16
 * leaving about 32 digits per bucket). */
int kinet_hip_table_window_bits(size_t n);
/* Device bytes of a table of n points (0 for a bad curve or c). */
size_t kinet_hip_table_bytes(uint32_t curve, size_t n, int window_bits);
/* Builds the table of n prepared points (d_inf may be NULL) into d_table
 * (kinet_hip_table_bytes).  Asynchronous on `stream`. */
int kinet_hip_points_table_build(uint32_t curve, const void* d_pts, const void* d_inf, size_t n, int window_bits,
                                 void* d_table, void* stream);
/* MSM of the first n <= table_n points of a table built for table_n points
 * with these window bits; d_inf = the points' identity flags (or NULL).
 * Output as kinet_hip_msm_prepared. */
int kinet_hip_msm_table(uint32_t curve, const void* d_scalars, const void* d_table, const void* d_inf,
                        size_t table_n, size_t n, int window_bits, void* d_out, void* stream);

/* Sets *d_bad (device int, zeroed by the caller) nonzero if any non-identity
 * prepared point is off the curve or, for BLS12-381, outside G1 (tested as
 * phi(P) = lambda P).  Curves 0 (secp256k1), 1, 2.  Asynchronous on
 * `stream`. */
int kinet_hip_points_check_g1(uint32_t curve, const void* d_pts, const void* d_inf, size_t n, int* d_bad,
                              void* stream);

/* Decompresses n G1 points (48-byte ZCash BLS12-381 / 32-byte BN254
 * encodings, the compressed form of the MSM outputs) into wire records
 * (big-endian x || y, all zero for infinity).  Sets *d_bad (device int, zeroed
 * by the caller) nonzero for any rejected encoding: missing compression flag,
 * infinity with other bits set, x >= p, or x^3 + b not a square (the rejected
 * record is written as zeros).  Subgroup membership is checked separately
 * (kinet_hip_points_check_g1).  No first-party counterpart in the reference
 * (it uses blst's codec, bls/cpp/bls_signature.cpp:58).  Asynchronous on
 * `stream`; gpukit codes. */
int kinet_hip_points_decompress(uint32_t curve, const void* d_comp, size_t n, void* d_wire, int* d_bad, void* stream);

/* ------------------------------------------------------------- batches
 * Many independent MSMs as one pipeline: every stage is launched once for
 * all jobs (bucket ids, windows and outputs of the jobs laid end to end), so
 * the per-MSM launch overhead and the serial window-combine chain (~256
 * doublings per MSM) are paid once per batch instead of once per job. */
typedef struct kinet_msm_batch kinet_msm_batch;

/* Plan for `count` MSMs over one prepared point array and one scalar array.
 * Job j: n[j] points from point pt_off[j] (pt_off NULL: every job starts at
 * point 0, e.g. prefixes of one SRS) and n[j] scalars from scalar sc_off[j]
 * (sc_off NULL: back to back in job order).  window_bits 0 = per job from the
 * installed tuning table (else the default rule), or 5..18 for every job.
 * The plan owns its scratch: different plans may run concurrently on
 * different streams; calls on one plan must not overlap.  gpukit codes. */
int kinet_hip_msm_batch_create(uint32_t curve, const size_t* n, const uint64_t* sc_off, const uint32_t* pt_off,
                               size_t count, int window_bits, kinet_msm_batch** out);
void kinet_hip_msm_batch_destroy(kinet_msm_batch* batch);
/* Window bits the plan uses for job `job` (-1 if out of range). */
int kinet_hip_msm_batch_window_bits(const kinet_msm_batch* batch, size_t job);
/* Runs every job; job j's record (wire point, then compressed: 96 + 48 or
 * 64 + 32 bytes) goes to d_out + j * record.  use_graph = 1 captures the
 * launch sequence in a HIP graph on first use with these pointers and this
 * stream and replays the graph afterwards (stream must not be NULL);
 * use_graph = 0 launches the kernels directly. */
int kinet_hip_msm_batch_run(kinet_msm_batch* batch, const void* d_scalars, const void* d_pts, const void* d_inf,
                            void* d_out, int use_graph, void* stream);

/* ------------------------------------------------------------- job sets
 * Many MSMs and Fr NTTs of uneven sizes as ONE HIP graph launch.  The MSMs
 * form one batch plan (as kinet_hip_msm_batch_*); the NTTs come in groups of
 * same-size transforms laid back to back.  The first run with a set of
 * pointers and a stream captures the whole launch sequence -- the NTT groups
 * on a second stream forked inside the capture, so they overlap the MSM
 * pipeline -- and later runs replay it with one hipGraphLaunch.  Outputs are
 * byte-identical to running the same MSMs (kinet_hip_msm_prepared) and NTTs
 * (kinet_hip_ntt_fr) one by one.  No counterpart in the reference. */
typedef struct kinet_job_set kinet_job_set;
typedef struct kinet_ntt_group {
    uint32_t log_n;   /* transform size 2^log_n */
    uint32_t count;   /* transforms in the group, 1..65535 */
    uint64_t offset;  /* first element of the group in the NTT in/out/tmp arrays */
} kinet_ntt_group;
/* MSM jobs as for kinet_hip_msm_batch_create (msm_count may be 0; curve 1
 * or 2; window bits from the tuning table); NTT groups over `field` (0
 * BLS12-381 Fr, 1 BN254 Fr; ngroups may be 0), all forward (ntt_inverse 0)
 * or all inverse.  Builds the MSM plan, its scratch and the twiddle tables
 * now.  gpukit codes. */
int kinet_hip_job_set_create(uint32_t curve, const size_t* msm_n, const uint64_t* sc_off, const uint32_t* pt_off,
                             size_t msm_count, uint32_t field, int ntt_inverse, const kinet_ntt_group* groups,
                             size_t ngroups, kinet_job_set** out);
void kinet_hip_job_set_destroy(kinet_job_set* set);
/* Window bits of MSM job `job` (-1 if out of range). */
int kinet_hip_job_set_window_bits(const kinet_job_set* set, size_t job);
/* Runs the set on `stream` (not NULL): MSM job j's record (wire point, then
 * compressed) at d_msm_out + j * record; NTT group g from d_ntt_in + 32 *
 * offset into d_ntt_out (d_ntt_tmp: scratch of the same layout).  *d_err is
 * cleared by the run itself and becomes nonzero if an NTT input is >= r.
 * Unused pointers may be NULL.  Runs of one set must not overlap. */
int kinet_hip_job_set_run(kinet_job_set* set, const void* d_scalars, const void* d_pts, const void* d_inf,
                          void* d_msm_out, const void* d_ntt_in, void* d_ntt_out, void* d_ntt_tmp, int* d_err,
                          void* stream);

/* ------------------------------------------------------------- tuning
 * Schedules are chosen per size class k = floor(log2 n) from in-process
 * tables; an empty cell means the built-in rule.  MSM cells are per curve
 * and mode (0: single MSM, 1: a job of a batch, 2: a single MSM over G1
 * points through the GLV split, kinet_hip_msm_prepared_g1) and hold the
 * window bits c and the first-level reduction segment; NTT cells are per
 * field and hold the pass split (log2 of the sub-transform lengths of passes
 * 1..3).  Schedules never change results, only speed.  A mode-2 cell of
 * window bits 1 records that the plain MSM was faster at that size class:
 * kinet_hip_msm_prepared_g1 called with window_bits 0 then runs it.
 *
 * Window bits only: c_by_log[k] for 2^k <= n < 2^(k+1), 0 = rule, count <= 32
 * (0 clears).  The rule is c = floor(log2 n) - 4 clamped to [5, 16] (of 2n
 * for mode 2). */
int kinet_hip_msm_set_window_table(uint32_t curve, int mode, const int* c_by_log, int count);
/* Window bits an n-point MSM of `mode` uses now (-1 for a bad curve / mode). */
int kinet_hip_msm_window_bits(uint32_t curve, size_t n, int mode);

/* Autotuner: times the candidate schedules of every size class 2^log_min ..
 * 2^log_max on the current device with synthetic inputs (multiples of the
 * generator, seeded scalars / field elements), checks that every candidate's
 * output equals the built-in schedule's byte for byte (GPUKIT_ERR_DETERMINISM
 * otherwise), and installs the fastest.  Candidates: MSM window bits around
 * log2 n - 4, then reduction segments 16/8/4/2 against the built-in rule for
 * the best c (modes 0, 2; batches of equal-size jobs totalling >= 2^20 points
 * for mode 1; for mode 2 also the plain MSM under its mode-0 cell); NTT every
 * split of log2 n into one to three passes of at most 2^10.  Synchronous;
 * holds the library's device lock, so other calls wait.  Mode 2 needs
 * log_min >= 11 (smaller GLV MSMs run the plain path).  Curve 0 (secp256k1)
 * has mode 0 only.  Sizes are limited to log2 n <= 24. */
int kinet_hip_tune_msm(uint32_t curve, int mode, int log_min, int log_max);
int kinet_hip_tune_ntt(uint32_t field, int log_min, int log_max);

/* The tuning tables as text, one cell per line ("msm <curve> <mode> <k> <c>
 * <seg>", "ntt <field> <k> <p1> <p2> <p3>", '#' comments): save writes every
 * set cell; load installs the cells of a file (validated; nothing is
 * installed if any line is malformed).  The engine loads the file named by
 * the environment variable KINET_HIP_TUNE_FILE, if set, on first use. */
int kinet_hip_tune_save(const char* path);
int kinet_hip_tune_load(const char* path);
/* Drops every tuned cell (back to the built-in rules). */
void kinet_hip_tune_clear(void);
/* Current schedule of a size class: MSM c and segment (0 = rule), NTT pass
 * split (p1, p2, p3; unused passes 0).  GPUKIT_ERR_BAD_SIZE for bad ids. */
int kinet_hip_tune_get_msm(uint32_t curve, int mode, int log_n, int* c, int* seg);
int kinet_hip_tune_get_ntt(uint32_t field, int log_n, int* p1, int* p2, int* p3);

/* Fixed-base batch scalar multiplication: out[i] = k_i * G for n scalars
 * (32-byte LE, any 256-bit value) on the device, written in the point wire
 * format (BE affine x || y, identity = zero bytes).  The table of multiples of
 * G is built on first use per curve and stays resident.  Synthetic MSM inputs
 * and SRS points [tau^i] G come from here.  Curves 0-3 (Banderwagon: the
 * Bandersnatch base point of banderwagon/cpp/element.cpp:48-75).
 * gpukit_status codes. */
int kinet_hip_g1_fixed_base_mul(uint32_t curve, const void* d_scalars, size_t n, void* d_out_wire, void* stream);

/* Opt-in stage profiler of the MSM pipeline: while enabled, each MSM records
 * HIP events on its own stream around its stages (0 digit sort, 1 bucket
 * accumulation, 2 bucket reduction, 3 window combine + output).  Collect sums
 * the elapsed milliseconds and call counts per stage (waiting for the events)
 * and forgets them. */
void kinet_hip_profile_enable(int on);
int kinet_hip_profile_collect(double* stage_ms, int* stage_calls, int nstages);

/* Tuning knob of the bucket reduction: buckets per first-level running-sum
 * segment (a power of two in [1, 16]; 0 restores the built-in rule) for MSMs
 * and batch plans created afterwards.  Results do not depend on it.  Returns
 * GPUKIT_ERR_BAD_SIZE for other values.  No counterpart in the reference
 * (its reduction is a CPU loop). */
int kinet_hip_msm_set_reduce_segment(int seg);

/* Montgomery batch inversion of device-resident base-field elements (curve
 * 0 secp256k1, 1 BN254: 32-byte LE records; 2 BLS12-381: 48-byte LE records),
 * asynchronous on `stream`: the device form of gpukit_batch_inv_<field>_fp_hip
 * (kinet/gpukit/batch_inversion.h), whose results it gives.  *d_status is
 * written when the work completes: 0, or 1 when an element is zero mod p
 * (outputs undefined; the host form returns GPUKIT_ERR_BAD_SIZE).  d_in and d_out
 * must not overlap and must be 4-byte aligned (GPUKIT_ERR_BAD_SIZE).  The
 * launches allocate nothing and may be captured in a graph.  No counterpart
 * in the reference (its GPU forms take host buffers). */
int kinet_hip_batch_inv_fp(uint32_t curve, const void* d_in, void* d_out, size_t n, uint32_t* d_status,
                           void* stream);

/* Measured VALU throughput of the current device, in instructions per second
 * summed over all lanes: op 0 v_mad_u64_u32 (the roofline denominator of the
 * MSM and NTT kernels), 1 v_mul_lo_u32 + add, 2 v_mul_hi_u32 + xor,
 * 3 v_mad_u32_u24, 4 v_fma_f64, 5 add3 + shift.  Returns 0 on success. */
int kinet_hip_valu_peak(int op, double* ops_per_second);

#ifdef __cplusplus
} /* extern "C" */
#endif
