// Device exclusive prefix sum over u32 counters (bucket offsets for the
// counting sort).  Three launches, no host round trip: block-local scans,
// one scan over the block totals, then a fix-up add.  The reference's CUDA
// scan (gpukit/gpu/cuda/prefix_sum.cu:24-72) scans block sums on the host;
// here everything stays on the device so the MSM pipeline can be captured
// in a HIP graph.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace kz {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 4;
constexpr int kScanTile = kScanThreads * kScanItems;

// Exclusive scan of n counters into out[0..n] (out[n] = total); also copies the
// offsets into cursor[] when it is not null.  block_sums needs
// scan_block_count(n) words, total one word.
void device_exclusive_scan(const uint32_t* d_in, uint32_t* d_out, uint32_t* d_cursor, uint32_t n,
                           uint32_t* d_block_sums, uint32_t* d_total, hipStream_t s);

inline uint32_t scan_block_count(uint32_t n) { return (n + kScanTile - 1) / kScanTile; }

// p[0..n) = v by a kernel (not hipMemsetAsync, so the clear is safe to record
// in a HIP graph and replay).
void device_fill_u32(uint32_t* d_p, uint32_t v, size_t n, hipStream_t s);

}  // namespace kz
