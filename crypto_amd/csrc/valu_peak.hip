// VALU throughput microbenchmark: the roofline denominator for the MSM and NTT
// kernels, which are bound by 32-bit integer multiply-adds, not by MFMA or HBM.
//
// Each lane runs 8 independent dependency chains of one instruction kind;
// 2048 waves per 256 CUs keep every SIMD busy.  Rate = lanes * iters * 8 / time.
#include <hip/hip_runtime.h>

#include "common.h"
#include "kinet/crypto/hip_runtime.h"

namespace kz {

constexpr int kPeakIters = 2048;
constexpr int kPeakChains = 8;

template <int OP>
__global__ void __launch_bounds__(256) k_valu_peak(uint64_t* __restrict__ out, uint32_t y, double fy) {
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t a[kPeakChains];
    uint32_t b[kPeakChains];
    double d[kPeakChains];
#pragma unroll
    for (int k = 0; k < kPeakChains; ++k) {
        a[k] = tid * 0x9E3779B97F4A7C15ull + k;
        b[k] = (uint32_t)a[k] | 1u;
        d[k] = (double)(tid + k);
    }
    for (int it = 0; it < kPeakIters; ++it) {
#pragma unroll
        for (int k = 0; k < kPeakChains; ++k) {
            if constexpr (OP == 0) a[k] = (uint64_t)(uint32_t)a[k] * y + a[k];           // v_mad_u64_u32
            if constexpr (OP == 1) b[k] = b[k] * b[k] + y;                               // v_mul_lo_u32 + add
            if constexpr (OP == 2) b[k] = __umulhi(b[k], y) ^ b[k];                      // v_mul_hi_u32
            if constexpr (OP == 3) b[k] = __umul24(b[k], y) + b[k];                      // v_mad_u32_u24
            if constexpr (OP == 4) d[k] = fma(d[k], fy, 1.0);                            // v_fma_f64
            if constexpr (OP == 5) b[k] = b[k] + y + (b[k] >> 1);                        // v_add3 class
        }
    }
    uint64_t acc = 0;
#pragma unroll
    for (int k = 0; k < kPeakChains; ++k) acc ^= a[k] ^ b[k] ^ (uint64_t)d[k];
    out[tid] = acc;
}

template <int OP>
static double run_peak(int repeats) {
    int dev = 0, cus = 0;
    KZ_HIP(hipGetDevice(&dev));
    KZ_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const unsigned blocks = (unsigned)cus * 16;  // 16 x 256 threads = 64 waves per CU
    const size_t threads = (size_t)blocks * 256;
    uint64_t* d_out = nullptr;
    KZ_HIP(hipMalloc(&d_out, threads * sizeof(uint64_t)));
    hipEvent_t e0, e1;
    KZ_HIP(hipEventCreate(&e0));
    KZ_HIP(hipEventCreate(&e1));
    hipLaunchKernelGGL(k_valu_peak<OP>, dim3(blocks), dim3(256), 0, 0, d_out, 0x9E3779B1u, 0.999999);
    KZ_HIP(hipEventRecord(e0, 0));
    for (int r = 0; r < repeats; ++r)
        hipLaunchKernelGGL(k_valu_peak<OP>, dim3(blocks), dim3(256), 0, 0, d_out, 0x9E3779B1u + r, 0.999999);
    KZ_HIP(hipEventRecord(e1, 0));
    KZ_HIP(hipEventSynchronize(e1));
    float ms = 0;
    KZ_HIP(hipEventElapsedTime(&ms, e0, e1));
    KZ_HIP(hipEventDestroy(e0));
    KZ_HIP(hipEventDestroy(e1));
    KZ_HIP(hipFree(d_out));
    return (double)threads * kPeakIters * kPeakChains * repeats / (ms * 1e-3);
}

}  // namespace kz

extern "C" int kinet_hip_valu_peak(int op, double* ops_per_second) {
    if (!ops_per_second) return -1;
    try {
        switch (op) {
            case 0: *ops_per_second = kz::run_peak<0>(5); break;
            case 1: *ops_per_second = kz::run_peak<1>(5); break;
            case 2: *ops_per_second = kz::run_peak<2>(5); break;
            case 3: *ops_per_second = kz::run_peak<3>(5); break;
            case 4: *ops_per_second = kz::run_peak<4>(5); break;
            case 5: *ops_per_second = kz::run_peak<5>(5); break;
            default: return -2;
        }
        return 0;
    } catch (const std::exception& ex) {
        kz::set_last_error(ex.what());
        return -4;
    }
}
