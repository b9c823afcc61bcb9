// BN254 G1: bucket-reduction level kernel (msm_impl.h).
#include "msm_impl.h"
#include "msm_stages.h"

namespace kz {

void launch_reduce_level_bn(unsigned grid, hipStream_t s, const uint32_t* inA, const uint32_t* inB,
                            const RedLevel* tab, uint32_t ntab, uint32_t uni_n, uint32_t uni_parts, uint32_t lm,
                            uint32_t* outA, uint32_t* outB) {
    hipLaunchKernelGGL(k_reduce_level<CurveBN254::Fp>, dim3(grid), dim3(kLevelThreads), 0, s, inA, inB, tab, ntab,
                       uni_n, uni_parts, lm, outA, outB);
}

}  // namespace kz
