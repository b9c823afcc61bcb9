// BLS12-381 G1: first-level bucket reduction kernel (msm_impl.h).
#include "msm_impl.h"
#include "msm_stages.h"
#include "scan.h"

namespace kz {

void launch_fold_bls(hipStream_t s, const uint32_t* offsets, uint32_t nb, uint32_t chunk, uint32_t nchunks,
                     uint32_t levels, uint32_t* list, uint32_t* runs, uint32_t* ctl, uint32_t* carry) {
    device_fill_u32(ctl, 0, 3, s);
    uint4* l4 = reinterpret_cast<uint4*>(list);
    uint2* r2 = reinterpret_cast<uint2*>(runs);
    hipLaunchKernelGGL(k_carry_runs<CurveBLS12381::Fp>, dim3(ceil_div((size_t)nb * 4, 256)), dim3(256), 0, s, offsets, nb,
                       chunk, r2, ctl, carry);
    hipLaunchKernelGGL(k_list_runs<CurveBLS12381::Fp>, dim3(256), dim3(256), 0, s, r2, l4, ctl);
    // levels = pairwise levels a run of this MSM can need; each group level
    // covers six of them.  Group leaders sit every 64^(level+1) list entries:
    // an odd grid stride spreads them over all workgroups (a multiple of 64
    // would hand every leader of a run to the same few).
    const unsigned grid = nchunks < 2047 ? (nchunks ? nchunks | 1u : 1u) : 2047;
    uint32_t stride = 1;
    for (uint32_t level = 0; level < levels; level += 6, stride *= kFoldFan)
        hipLaunchKernelGGL(k_fold_groups<CurveBLS12381::Fp>, dim3(grid), dim3(4 * kFoldFan), 0, s, l4, ctl, stride, carry);
}

void launch_reduce_seg_bls(unsigned grid, hipStream_t s, const uint32_t* offsets, uint32_t chunk,
                            const uint32_t* bsum, const uint32_t* carry, uint32_t nseg, uint32_t seg, uint32_t* segA,
                            uint32_t* segB) {
    hipLaunchKernelGGL(k_reduce_seg<CurveBLS12381::Fp>, dim3(grid), dim3(256), 0, s, offsets, chunk, bsum, carry, nseg, seg, segA,
                       segB);
}

}  // namespace kz
