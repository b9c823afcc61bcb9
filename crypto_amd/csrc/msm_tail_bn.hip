// BN254 G1: window-table rows and compression kernels (msm_impl.h); the
// window combine is in msm_combine_row.hip.
#include "msm_impl.h"
#include "msm_stages.h"

namespace kz {

void launch_table_rows_bn(const uint32_t* pts, const uint8_t* inf, uint32_t n, uint32_t c, uint32_t nwin,
                         uint32_t* table, hipStream_t s) {
    hipLaunchKernelGGL(k_table_rows<CurveBN254>, dim3(ceil_div(n, 64)), dim3(64), 0, s, pts, inf, n, c, nwin, table);
}

void launch_compress_bn(const uint8_t* wire, uint8_t* out, hipStream_t s) {
    hipLaunchKernelGGL(k_compress_point<CurveBN254>, dim3(1), dim3(64), 0, s, wire, out);
}

}  // namespace kz
