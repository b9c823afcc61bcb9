// Host interface of the scalar-field NTT (BLS12-381 Fr, BN254 Fr).
//
// Domain: X_k = sum_j x_j w^(jk), w the primitive n-th root g^((r-1)/n) with
// g = 7 (BLS12-381 Fr) or 5 (BN254 Fr); natural order in and out; the inverse
// scales by n^-1.  Elements are 32-byte little-endian canonical integers, in
// and out, which is also the MSM scalar wire format.
#pragma once

#include "common.h"

namespace kz {

constexpr int kNttMaxLog = 26;      // largest supported transform: 2^26 elements
constexpr int kNttPassMaxLog = 11;  // longest sub-transform of one pass (its LDS tile)

// Forward (inverse = false) or inverse transform of n = 2^logn elements, from
// d_in to d_out (may alias) through the scratch buffer d_tmp (n * 32 bytes).
// d_err (one int, zeroed by the caller) gets nonzero if an input is >= r.
void ntt_device(int field, int logn, bool inverse, const uint32_t* d_in, uint32_t* d_out, uint32_t* d_tmp,
                int* d_err, hipStream_t s);

// count transforms of the same size laid back to back (d_in, d_out, d_tmp:
// count * n * 32 bytes), one launch per pass for all of them.
void ntt_device_batch(int field, int logn, bool inverse, const uint32_t* d_in, uint32_t* d_out, uint32_t* d_tmp,
                      size_t count, int* d_err, hipStream_t s);

// Builds the twiddle tables of (field, 2^logn, direction) under the split in
// effect now (allocates and synchronises), so later transforms only launch
// kernels.
void ntt_prepare(int field, int logn, bool inverse, hipStream_t s);

// Four-step tables built after a mark are freed by ntt_drop_tables_since (the
// tuner's trials; the stream must be idle).  Tables built before the mark,
// which captured graphs may hold, stay.
size_t ntt_table_mark();
void ntt_drop_tables_since(size_t mark);

// Pass split of a size class: log2 of the sub-transform lengths of passes
// 1..3 (each <= 11, summing to logn, unused passes 0 at the end).  The tuned
// cell (all zero = built-in split) and the split in effect.
bool ntt_split_valid(int logn, int p1, int p2, int p3);
void ntt_set_schedule(int field, int logn, int p1, int p2, int p3);
void ntt_get_schedule(int field, int logn, int& p1, int& p2, int& p3);
void ntt_schedule_of(int field, int logn, int& p1, int& p2, int& p3);

}  // namespace kz
