// BN254 G1: wire conversion and bucket accumulation kernels (msm_impl.h).
#include "msm_impl.h"
#include "msm_stages.h"

namespace kz {

void launch_points_from_wire_bn(const uint8_t* wire, uint32_t* pts, uint8_t* inf, uint32_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_points_from_wire<CurveBN254>, dim3(ceil_div(n, 128)), dim3(128), 0, s, wire, pts, inf, n);
}

void launch_accumulate_bn(unsigned grid, hipStream_t s, const uint32_t* entries, const uint32_t* offsets,
                            uint32_t nb, uint32_t chunk, const uint32_t* pts, uint32_t* bsum, uint32_t* carry) {
    hipLaunchKernelGGL(k_accumulate<CurveBN254::Fp>, dim3(grid), dim3(64), 0, s, entries, offsets, nb, chunk, pts, bsum,
                       carry);
}

}  // namespace kz
