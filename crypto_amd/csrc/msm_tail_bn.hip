// BN254 G1: window combine, normalisation and compression kernels (msm_impl.h).
#include "msm_impl.h"
#include "msm_stages.h"

namespace kz {

void launch_combine_bn(hipStream_t s, const uint32_t* winres, const MsmJob* jobs, const MsmJob& one,
                       uint32_t njobs, uint8_t* out) {
    hipLaunchKernelGGL(k_combine<CurveBN254>, dim3(ceil_div((size_t)njobs * 4, 64)), dim3(64), 0, s, winres, jobs,
                       one, njobs, out);
}

void launch_table_rows_bn(const uint32_t* pts, const uint8_t* inf, uint32_t n, uint32_t c, uint32_t nwin,
                         uint32_t* table, hipStream_t s) {
    hipLaunchKernelGGL(k_table_rows<CurveBN254>, dim3(ceil_div(n, 64)), dim3(64), 0, s, pts, inf, n, c, nwin, table);
}

void launch_compress_bn(const uint8_t* wire, uint8_t* out, hipStream_t s) {
    hipLaunchKernelGGL(k_compress_point<CurveBN254>, dim3(1), dim3(64), 0, s, wire, out);
}

}  // namespace kz
