// BN254 G1 instantiation of the MSM pipeline (kernels in msm_impl.h).
#include "msm_impl.h"

namespace kz {

void msm_device_bn(const uint32_t* d_pts, const uint8_t* d_inf, const uint32_t* d_scalars, size_t n,
                    const MsmParams& prm, MsmWorkspace& ws, hipStream_t s, uint8_t* d_out) {
    msm_device_t<CurveBN254>(d_pts, d_inf, d_scalars, n, prm, ws, s, d_out);
}

void points_from_wire_bn(const uint8_t* d_wire, uint32_t* d_pts, uint8_t* d_inf, size_t n, hipStream_t s) {
    points_from_wire_t<CurveBN254>(d_wire, d_pts, d_inf, n, s);
}

void compress_bn(const uint8_t* d_wire, uint8_t* d_out, hipStream_t s) { compress_t<CurveBN254>(d_wire, d_out, s); }

}  // namespace kz
