// BLS12-381 G1: window combine, normalisation and compression kernels (msm_impl.h).
#include "msm_impl.h"
#include "msm_stages.h"

namespace kz {

void launch_combine_bls(hipStream_t s, const uint32_t* winres, int nwin, int c, uint8_t* out) {
    hipLaunchKernelGGL(k_combine<CurveBLS12381>, dim3(1), dim3(64), 0, s, winres, nwin, c, out);
}

void launch_compress_bls(const uint8_t* wire, uint8_t* out, hipStream_t s) {
    hipLaunchKernelGGL(k_compress_point<CurveBLS12381>, dim3(1), dim3(64), 0, s, wire, out);
}

}  // namespace kz
