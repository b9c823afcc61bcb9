// secp256k1: bucket-reduction levels and window-table rows (msm_impl.h) for
// the multi-curve entry; the window combine is in msm_combine_row.hip.
#include "msm_impl.h"
#include "msm_stages.h"

namespace kz {

using Secp = CurveSecp256k1;

void launch_reduce_level_secp(unsigned grid, hipStream_t s, const uint32_t* inA, const uint32_t* inB,
                              const RedLevel* tab, uint32_t ntab, uint32_t uni_n, uint32_t uni_parts, uint32_t lm,
                              uint32_t* outA, uint32_t* outB) {
    hipLaunchKernelGGL(k_reduce_level<Secp::Fp>, dim3(grid), dim3(kLevelThreads), 0, s, inA, inB, tab, ntab, uni_n,
                       uni_parts, lm, outA, outB);
}

void launch_table_rows_secp(const uint32_t* pts, const uint8_t* inf, uint32_t n, uint32_t c, uint32_t nwin,
                            uint32_t* table, hipStream_t s) {
    hipLaunchKernelGGL(k_table_rows<Secp>, dim3(ceil_div(n, 64)), dim3(64), 0, s, pts, inf, n, c, nwin, table);
}

// secp256k1 records carry no compressed form: nothing to launch.
void launch_compress_secp(const uint8_t*, uint8_t*, hipStream_t) {}

}  // namespace kz
