// G1 point decompression on the device: the compressed encodings that the MSM
// outputs carry (write_compressed in msm_impl.h) back to the wire layout.
//
//   BLS12-381, 48 bytes (ZCash): bit 7 of byte 0 set (compressed); bit 6 =
//     infinity, then byte 0 is 0xC0 and every other byte zero; bit 5 = y is
//     the larger of y, p - y; x = the remaining 381 bits, below p.
//   BN254, 32 bytes: top two bits 0b10 (smaller y) / 0b11 (larger y) / 0b01
//     (infinity, byte 0 = 0x40, the rest zero); x = the remaining 254 bits.
// y = (x^3 + b)^((p+1)/4) (p = 3 mod 4 on both curves), accepted only when
// its square is x^3 + b.  Output: wire records (big-endian x || y, all zero
// for infinity); any rejected encoding sets *bad.  Subgroup membership is the
// separate kinet_hip_points_check_g1.  There is no first-party codec in the
// reference (blst's, bls/cpp/bls_signature.cpp:58); the oracle's
// orc_g1_decompress states the same rules on the CPU.
#include "curve.h"
#include "engine.h"
#include "kinet/crypto/hip_runtime.h"
#include "kinet/gpukit/gpukit.h"

namespace kz {
namespace {

// (p + 1) / 4 as 32-bit LE words
constexpr uint32_t kSqrtExpBls[12] = {0xFFFFEAABu, 0xEE7FBFFFu, 0xAC54FFFFu, 0x07AAFFFFu, 0x3DAC3D89u, 0xD9CC34A8u,
                                      0x3CE144AFu, 0xD91DD2E1u, 0x90D2EB35u, 0x92C6E9EDu, 0x8E5FF9A6u, 0x0680447Au};
constexpr uint32_t kSqrtExpBn[8] = {0xB61F3F52u, 0x4F082305u, 0x5A1C72A3u, 0x65E05AA4u,
                                    0xA0605617u, 0x6E14116Du, 0xB84C680Au, 0x0C19139Cu};

// a^e for the exponent words e (most significant bit first), Montgomery form
template <class F, int W>
__device__ Fe<F> fe_pow_words(const Fe<F>& a, const uint32_t (&e)[W]) {
    Fe<F> r = fe_const<F>(F::ONE);
#pragma unroll 1
    for (int bit = 32 * W - 1; bit >= 0; --bit) {
        r = fe_sqr<F>(r);
        if ((e[bit >> 5] >> (bit & 31)) & 1u) r = fe_mul<F>(r, a);
    }
    return r;
}

template <class C>
__global__ void __launch_bounds__(64)
k_decompress(const uint8_t* __restrict__ comp, uint32_t n, uint8_t* __restrict__ wire, int* __restrict__ bad) {
    using F = typename C::Fp;
    constexpr int FB = C::FB;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint8_t* c = comp + (size_t)i * FB;
    uint8_t* out = wire + (size_t)i * 2 * FB;
    uint8_t xb[FB];
    uint32_t rest = 0;
    for (int k = 0; k < FB; ++k) {
        xb[k] = c[k];
        if (k) rest |= c[k];
    }
    const uint8_t f = c[0];
    bool ok, larger = false, inf = false;
    if (C::ID == 2) {
        ok = (f & 0x80) != 0;
        if (ok && (f & 0x40)) {
            inf = true;
            ok = f == 0xC0 && rest == 0;
        }
        larger = (f & 0x20) != 0;
        xb[0] &= 0x1F;
    } else {
        const int flag = f >> 6;
        ok = flag != 0;
        if (flag == 1) {
            inf = true;
            ok = f == 0x40 && rest == 0;
        }
        larger = flag == 3;
        xb[0] &= 0x3F;
    }
    if (ok && inf) {
        for (int k = 0; k < 2 * FB; ++k) out[k] = 0;
        return;
    }
    Fe<F> y_canon = fe_zero<F>();
    Fe<F> x = fe_from_be_bytes<F>(xb, FB);
    if (ok) ok = fe_is_canonical<F>(x);
    if (ok) {
        const Fe<F> xm = fe_to_mont<F>(x);
        Fe<F> b = fe_zero<F>();
        b.v[0] = C::B;
        const Fe<F> rhs = fe_add<F>(fe_mul<F>(fe_sqr<F>(xm), xm), fe_to_mont<F>(b));
        Fe<F> y;
        if constexpr (C::ID == 2) y = fe_pow_words<F>(rhs, kSqrtExpBls);
        else y = fe_pow_words<F>(rhs, kSqrtExpBn);
        ok = fe_eq<F>(fe_sqr<F>(y), rhs);  // rhs a square
        y_canon = fe_from_mont<F>(y);
        const Fe<F> ny = fe_neg<F>(y_canon);
        if (fe_gt<F>(y_canon, ny) != larger) y_canon = ny;
    }
    if (!ok) {
        atomicOr(bad, 1);
        for (int k = 0; k < 2 * FB; ++k) out[k] = 0;
        return;
    }
    for (int k = 0; k < FB; ++k) out[k] = xb[k];
    fe_to_be_bytes<F>(y_canon, out + FB, FB);
}

}  // namespace

void points_decompress(int curve, const uint8_t* d_comp, size_t n, uint8_t* d_wire, int* d_bad, hipStream_t s) {
    if (n == 0) return;
    const unsigned grid = ceil_div(n, 64);
    if (curve == 2)
        hipLaunchKernelGGL(k_decompress<CurveBLS12381>, dim3(grid), dim3(64), 0, s, d_comp, (uint32_t)n, d_wire, d_bad);
    else if (curve == 1)
        hipLaunchKernelGGL(k_decompress<CurveBN254>, dim3(grid), dim3(64), 0, s, d_comp, (uint32_t)n, d_wire, d_bad);
    else
        throw std::invalid_argument("curve");
    KZ_HIP(hipGetLastError());
}

}  // namespace kz

using namespace kz;

extern "C" int kinet_hip_points_decompress(uint32_t curve, const void* d_comp, size_t n, void* d_wire, int* d_bad,
                                           void* stream) {
    if (curve != 1 && curve != 2) return GPUKIT_ERR_BACKEND;
    if (!d_bad || (n > 0 && (!d_comp || !d_wire))) return GPUKIT_ERR_NULL_ARG;
    if (n >= (size_t)1 << 31) return GPUKIT_ERR_BAD_SIZE;
    try {
        (void)engine();
        points_decompress((int)curve, static_cast<const uint8_t*>(d_comp), n, static_cast<uint8_t*>(d_wire), d_bad,
                          static_cast<hipStream_t>(stream));
        return GPUKIT_OK;
    } catch (const std::exception& ex) {
        set_last_error(ex.what());
        return GPUKIT_ERR_BACKEND;
    }
}
