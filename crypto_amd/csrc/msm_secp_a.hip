// secp256k1: wire conversion, bucket accumulation and first-level reduction
// kernels (msm_impl.h) for the plain MSM of the multi-curve entry.
#include "msm_impl.h"
#include "msm_stages.h"
#include "scan.h"

namespace kz {

using Secp = CurveSecp256k1;

void launch_points_from_wire_secp(const uint8_t* wire, uint32_t* pts, uint8_t* inf, uint32_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_points_from_wire<Secp>, dim3(ceil_div(n, 128)), dim3(128), 0, s, wire, pts, inf, n);
}

void launch_accumulate_secp(unsigned grid, hipStream_t s, const uint32_t* entries, const uint32_t* offsets,
                            uint32_t nb, uint32_t chunk, const uint32_t* pts, uint32_t* bsum, uint32_t* carry) {
    hipLaunchKernelGGL(k_accumulate<Secp::Fp>, dim3(grid), dim3(64), 0, s, entries, offsets, nb, chunk, pts, bsum,
                       carry);
}

void launch_fold_secp(hipStream_t s, const uint32_t* offsets, uint32_t nb, uint32_t chunk, uint32_t nchunks,
                      uint32_t levels, uint32_t* list, uint32_t* runs, uint32_t* ctl, uint32_t* carry) {
    device_fill_u32(ctl, 0, 3, s);
    uint4* l4 = reinterpret_cast<uint4*>(list);
    uint2* r2 = reinterpret_cast<uint2*>(runs);
    hipLaunchKernelGGL(k_carry_runs<Secp::Fp>, dim3(ceil_div((size_t)nb * 4, 256)), dim3(256), 0, s, offsets, nb,
                       chunk, r2, ctl, carry);
    hipLaunchKernelGGL(k_list_runs<Secp::Fp>, dim3(256), dim3(256), 0, s, r2, l4, ctl);
    // Odd grid stride over the group leaders, as for the other curves.
    const unsigned grid = nchunks < 2047 ? (nchunks ? nchunks | 1u : 1u) : 2047;
    uint32_t stride = 1;
    for (uint32_t level = 0; level < levels; level += 6, stride *= kFoldFan)
        hipLaunchKernelGGL(k_fold_groups<Secp::Fp>, dim3(grid), dim3(4 * kFoldFan), 0, s, l4, ctl, stride, carry);
}

void launch_reduce_seg_secp(unsigned grid, hipStream_t s, const uint32_t* offsets, uint32_t chunk,
                            const uint32_t* bsum, const uint32_t* carry, uint32_t nseg, uint32_t seg, uint32_t* segA,
                            uint32_t* segB) {
    hipLaunchKernelGGL(k_reduce_seg<Secp::Fp>, dim3(grid), dim3(256), 0, s, offsets, chunk, bsum, carry, nseg, seg,
                       segA, segB);
}

}  // namespace kz
