// BN254 G1: first-level bucket reduction kernel (msm_impl.h).
#include "msm_impl.h"
#include "msm_stages.h"

namespace kz {

void launch_reduce_seg_bn(unsigned grid, hipStream_t s, const uint32_t* offsets, uint32_t chunk,
                            const uint32_t* bsum, const uint32_t* carry, uint32_t nseg, uint32_t* segA,
                            uint32_t* segB) {
    hipLaunchKernelGGL(k_reduce_seg<CurveBN254::Fp>, dim3(grid), dim3(64), 0, s, offsets, chunk, bsum, carry, nseg, segA,
                       segB);
}

}  // namespace kz
