// Internal host-side declarations shared by the MSM translation units.
#pragma once

#include <stdexcept>

#include "msm.h"

namespace kz {

// Signed-digit count (scatter = false) or scatter (scatter = true) pass.
// table_stride > 0: the points are a window table of that row stride, and
// every window's digits share one bucket set (keys |d| - 1, entries w T + i).
void digits_dispatch(int c, bool scatter, const uint32_t* d_scalars, const uint8_t* d_inf, uint32_t n,
                     uint32_t table_stride, uint32_t* counts_or_cursor, uint32_t* entries, hipStream_t s);

// LDS-histogram sort of one MSM (c in [5, 16], n >= 4096): fills offsets
// (W 2^(c-1) + 1 words) and entries.  cnt and pre need lds_sort_words(c, n)
// words each, tot W 2^(c-1) words.
bool lds_sort_supported(int c, size_t n);
void lds_sort_shape(int c, int bits, size_t n, uint32_t& wpad, uint32_t& nparts, uint32_t& per_part);
inline size_t lds_sort_words(int c, int bits, size_t n) {
    uint32_t wpad, nparts, per_part;
    lds_sort_shape(c, bits, n, wpad, nparts, per_part);
    return (size_t)msm_num_windows(c, bits) * nparts * ((size_t)1 << (c - 1));
}
void lds_sort_dispatch(int c, int bits, const uint32_t* d_scalars, const uint8_t* d_inf, uint32_t n,
                       uint32_t table_stride, uint32_t* cnt, uint32_t* pre, uint32_t* tot, uint32_t* offsets,
                       uint32_t* block_sums, uint32_t* total, uint32_t* entries, hipStream_t s);

// The same passes over every job of a batch (d_jobs: njobs descriptors on the
// device, nblocks = total 256-thread blocks).
void digits_multi_dispatch(bool scatter, const uint32_t* d_scalars, const uint8_t* d_inf, const MsmJob* d_jobs,
                           uint32_t njobs, uint32_t nblocks, uint32_t* counts_or_cursor, uint32_t* entries,
                           hipStream_t s);

// Batch LDS-histogram sort (plan.lds_sort): digit codes, per-item counts,
// column scan (tot = the bucket counts, nb words), scan into offsets, scatter.
// dyn_lds = largest 2^(c-1) words of the batch, in bytes.
void batch_lds_sort_dispatch(const uint32_t* d_scalars, const uint8_t* d_inf, const MsmJob* d_jobs,
                             const SortJob* d_sjobs, uint32_t njobs, uint32_t dig_blocks, uint32_t items, uint32_t nb,
                             size_t dyn_lds, uint16_t* dig, uint32_t* cnt, uint32_t* pre, uint32_t* tot,
                             uint32_t* offsets, uint32_t* block_sums, uint32_t* total, uint32_t* entries, hipStream_t s);
// Allows the batch sort kernels dyn_lds bytes of dynamic LDS; false if the
// runtime refuses (the batch then sorts with global atomics).
bool batch_lds_sort_prepare(size_t dyn_lds);

}  // namespace kz
