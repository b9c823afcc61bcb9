// BLS12-381 G1: per-window bucket reduction kernel (msm_impl.h).
#include "msm_impl.h"
#include "msm_stages.h"

namespace kz {

void launch_reduce_window_bls(unsigned nwin, hipStream_t s, const uint32_t* segA, const uint32_t* segB,
                              uint32_t t1, const uint2* win_seg, uint32_t* winres) {
    hipLaunchKernelGGL(k_reduce_window<CurveBLS12381::Fp>, dim3(nwin), dim3(kReduceThreads), 0, s, segA, segB, t1,
                       win_seg, winres);
}

}  // namespace kz
