// Device exclusive prefix sum kernels (see scan.h).
#include "scan.h"

namespace kz {

// Exclusive scan of the 256 per-thread values of a block; returns the block total.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* sh, uint32_t& total) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int s = 1; s < kScanThreads; s <<= 1) {
        uint32_t add = (t >= s) ? sh[t - s] : 0u;
        __syncthreads();
        sh[t] += add;
        __syncthreads();
    }
    total = sh[kScanThreads - 1];
    return sh[t] - v;
}

__global__ void __launch_bounds__(kScanThreads)
k_scan_local(const uint32_t* __restrict__ in, uint32_t* __restrict__ out, uint32_t* __restrict__ block_sums,
             uint32_t n) {
    __shared__ uint32_t sh[kScanThreads];
    const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint32_t sum = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        v[i] = (base + i < n) ? in[base + i] : 0u;
        sum += v[i];
    }
    uint32_t total;
    uint32_t run = block_exclusive_scan(sum, sh, total);
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        if (base + i < n) out[base + i] = run;
        run += v[i];
    }
    if (threadIdx.x == 0) block_sums[blockIdx.x] = total;
}

// One block: exclusive scan of nb block totals in place; writes the grand total.
__global__ void __launch_bounds__(kScanThreads)
k_scan_blocks(uint32_t* __restrict__ block_sums, uint32_t nb, uint32_t* __restrict__ grand_total) {
    __shared__ uint32_t sh[kScanThreads];
    uint32_t carry = 0;
    for (uint32_t base = 0; base < nb; base += kScanThreads) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t v = (i < nb) ? block_sums[i] : 0u;
        uint32_t total;
        const uint32_t ex = block_exclusive_scan(v, sh, total);
        if (i < nb) block_sums[i] = carry + ex;
        carry += total;
        __syncthreads();
    }
    if (threadIdx.x == 0) *grand_total = carry;
}

// out[i] += block offset; also copies into cursor[] and writes out[n] = total.
__global__ void __launch_bounds__(kScanThreads)
k_scan_add(uint32_t* __restrict__ out, uint32_t* __restrict__ cursor, const uint32_t* __restrict__ block_sums,
           uint32_t n, const uint32_t* __restrict__ grand_total) {
    const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    const uint32_t off = block_sums[blockIdx.x];
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        if (base + i < n) {
            const uint32_t v = out[base + i] + off;
            out[base + i] = v;
            if (cursor) cursor[base + i] = v;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = *grand_total;
}

void device_exclusive_scan(const uint32_t* d_in, uint32_t* d_out, uint32_t* d_cursor, uint32_t n,
                                  uint32_t* d_block_sums, uint32_t* d_total, hipStream_t s) {
    const uint32_t nb = (n + kScanTile - 1) / kScanTile;
    hipLaunchKernelGGL(k_scan_local, dim3(nb), dim3(kScanThreads), 0, s, d_in, d_out, d_block_sums, n);
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(kScanThreads), 0, s, d_block_sums, nb, d_total);
    hipLaunchKernelGGL(k_scan_add, dim3(nb), dim3(kScanThreads), 0, s, d_out, d_cursor, d_block_sums, n,
                       d_total);
}

__global__ void __launch_bounds__(256) k_fill_u32(uint32_t* __restrict__ p, uint32_t v, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

void device_fill_u32(uint32_t* d_p, uint32_t v, size_t n, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, d_p, v, n);
}

}  // namespace kz
