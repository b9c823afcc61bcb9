// GLV endomorphism for BLS12-381 G1 MSMs, and the G1 membership check that
// licenses it.
//
// BLS12-381 G1 has the endomorphism phi(x, y) = (beta x, y), beta a cube root
// of unity in Fp, acting on G1 as multiplication by lambda =
// 0xac45a4010001a40200000000ffffffff (a cube root of unity mod r, 128 bits;
// the pair (beta, lambda) below was matched by checking phi(G) = lambda G).
// For P in G1 and any 256-bit s, with s = k1 + lambda k2 mod r (the signed
// split of glv.h, |k1|, |k2| < 2^127):
//     s P = |k1| (+-P) + |k2| (+-phi(P)),
// so an n-point MSM with 256-bit scalars becomes a 2n-point MSM with 128-bit
// scalars: half the windows, hence half the bucket reduction and half the
// Horner doublings of the window combine, for the same number of additions.
// The identity holds only for points of G1 (on E(Fp) outside G1, phi is not
// multiplication by lambda), so the GLV path is used only for point sets that
// passed k_check_g1: on the curve, and phi(P) = lambda P, which for BLS12
// curves holds exactly for the points of G1 (Scott, "A note on group
// membership tests for G1, G2 and GT on BLS pairing-friendly curves", 2021).
// BN254 G1 and secp256k1 have cofactor 1: their check is the curve equation
// alone.  (secp256k1 is curve 0 of the multi-curve entry; its split is
// glvsecp's, with magnitudes below 2^128.)
#include <type_traits>

#include "curve.h"
#include "engine.h"
#include "glv.h"
#include "kinet/crypto/hip_runtime.h"
#include "kinet/gpukit/gpukit.h"

namespace kz {

namespace {

// BN254: beta R mod p (radix 2^29, R = 2^261) for
// beta = 0x30644e72e131a0295e6dd9e7e0acccb0c28f069fbb966e3de4bd44e5607cfd48
constexpr uint32_t kBetaMontBn[9] = {0x18CCB791u, 0x175B1C3Au, 0x0B83D6E2u, 0x0E8ED071u, 0x1282BEE2u,
                                     0x04220E84u, 0x1FE4017Fu, 0x15084D4Au, 0x00169119u};
// secp256k1: beta R mod p (radix 2^29, R = 2^290) for
// beta = 0x7ae96a2b657c07106e64479eac3434e99cf0497512f58995c1396c28719501ee
constexpr uint32_t kBetaMontSecp[10] = {0x14EABFDAu, 0x197839E6u, 0x04361D08u, 0x1C057AC5u, 0x1858C712u,
                                        0x120A07FBu, 0x178D02E3u, 0x09FC74C4u, 0x00F2ECF5u, 0x00000000u};
// BLS12-381: the non-adjacent form of lambda = 0xac45a4010001a40200000000ffffffff
// as two digit masks, lambda = pos - neg (bits 0..128)
constexpr uint32_t kLamNafPos[5] = {0x00000000u, 0x00000001u, 0x00022402u, 0x00482401u, 0x00000001u};
constexpr uint32_t kLamNafNeg[5] = {0x00000001u, 0x00000000u, 0x00008000u, 0x54028000u, 0x00000000u};
// BLS12-381: beta R mod p (radix 2^29, R = 2^406)
constexpr uint32_t kBetaMont[14] = {0x1195DFEBu, 0x1B04E484u, 0x06026044u, 0x086070A2u, 0x1FD68858u,
                                    0x137E9670u, 0x06871E67u, 0x1E736664u, 0x083B24F6u, 0x08A70373u,
                                    0x02A012FDu, 0x0112F94Bu, 0x18A2733Cu, 0x00000003u};
}  // namespace

// Signed GLV split (glv.h): out_pts = [+-P_i, +-phi(P_i)], out_inf likewise,
// out_sc = [|k1_i|, |k2_i|] (16-byte records).
template <int CURVE>
__global__ void __launch_bounds__(256)
k_glv_expand(const uint32_t* __restrict__ pts, const uint8_t* __restrict__ inf, const uint32_t* __restrict__ scalars,
             uint32_t n, uint32_t* __restrict__ out_pts, uint8_t* __restrict__ out_inf, uint32_t* __restrict__ out_sc) {
    using F = std::conditional_t<CURVE == 0, SecpFp, std::conditional_t<CURVE == 1, BnFp, BlsFp>>;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t s[8];
    {
        const uint4* q = reinterpret_cast<const uint4*>(scalars + (size_t)i * 8);
        const uint4 a = q[0], b = q[1];
        s[0] = a.x; s[1] = a.y; s[2] = a.z; s[3] = a.w;
        s[4] = b.x; s[5] = b.y; s[6] = b.z; s[7] = b.w;
    }
    uint32_t k1[4], k2[4];
    bool neg1, neg2;
    if constexpr (CURVE == 0) glvsecp::split(s, k1, k2, neg1, neg2);
    else if constexpr (CURVE == 1) glvbn::split(s, k1, k2, neg1, neg2);
    else glvbls::split(s, k1, k2, neg1, neg2);
    Affine<F> a = affine_load<F>(pts, i);
    const Fe<F> y = a.y, ny = fe_neg<F>(a.y);
    a.y = neg1 ? ny : y;
    affine_store<F>(out_pts, i, a);
    if constexpr (CURVE == 0) a.x = fe_mul<F>(a.x, fe_const<F>(kBetaMontSecp));
    else if constexpr (CURVE == 1) a.x = fe_mul<F>(a.x, fe_const<F>(kBetaMontBn));
    else a.x = fe_mul<F>(a.x, fe_const<F>(kBetaMont));
    a.y = neg2 ? ny : y;
    affine_store<F>(out_pts, (size_t)n + i, a);
    const uint8_t f = inf ? inf[i] : 0;
    out_inf[i] = f;
    out_inf[(size_t)n + i] = f;
    uint4* o = reinterpret_cast<uint4*>(out_sc);
    o[i] = make_uint4(k1[0], k1[1], k1[2], k1[3]);
    o[(size_t)n + i] = make_uint4(k2[0], k2[1], k2[2], k2[3]);
}

// *bad |= 1 if a non-identity point is off the curve or (BLS12-381) outside G1.
template <class C>
__global__ void __launch_bounds__(64)
k_check_g1(const uint32_t* __restrict__ pts, const uint8_t* __restrict__ inf, uint32_t n, int* __restrict__ bad) {
    using F = typename C::Fp;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || (inf && inf[i])) return;
    const Affine<F> a = affine_load<F>(pts, i);
    Fe<F> b = fe_zero<F>();
    b.v[0] = C::B;
    const Fe<F> rhs = fe_add<F>(fe_mul<F>(fe_mul<F>(a.x, a.x), a.x), fe_to_mont<F>(b));
    bool ok = fe_eq<F>(fe_mul<F>(a.y, a.y), rhs);
    if constexpr (C::ID == 2) {
        if (!ok) {
            atomicOr(bad, 1);
            return;
        }
        // lambda P by signed-digit double-and-add over lambda's NAF (top digit
        // +1 at bit 128): 18 additions instead of lambda's 48 set bits
        // (-P is formed at each negative digit: kept live, its 2N registers
        // cost the kernel a wave per SIMD)
        Xyzz<F> acc = xyzz_from_affine<F>(a);
#pragma unroll 1
        for (int bit = 127; bit >= 0; --bit) {
            acc = xyzz_dbl<F, MulCall>(acc);
            if ((kLamNafPos[bit >> 5] >> (bit & 31)) & 1u) {
                xyzz_madd<F, MulCall>(acc, a);
            } else if ((kLamNafNeg[bit >> 5] >> (bit & 31)) & 1u) {
                Affine<F> na = a;
                na.y = fe_neg<F>(a.y);
                xyzz_madd<F, MulCall>(acc, na);
            }
        }
        const Fe<F> bx = fe_mul<F>(a.x, fe_const<F>(kBetaMont));
        ok = !xyzz_is_inf<F>(acc) && fe_eq<F>(acc.X, fe_mul<F>(bx, acc.ZZ)) && fe_eq<F>(acc.Y, fe_mul<F>(a.y, acc.ZZZ));
    }
    if (!ok) atomicOr(bad, 1);
}

void glv_expand(int curve, const uint32_t* d_pts, const uint8_t* d_inf, const uint32_t* d_scalars, size_t n,
                uint32_t* d_pts2, uint8_t* d_inf2, uint32_t* d_scalars2, hipStream_t s) {
    const unsigned grid = ceil_div(n, 256);
    if (curve == 0)
        hipLaunchKernelGGL(k_glv_expand<0>, dim3(grid), dim3(256), 0, s, d_pts, d_inf, d_scalars, (uint32_t)n, d_pts2,
                           d_inf2, d_scalars2);
    else if (curve == 1)
        hipLaunchKernelGGL(k_glv_expand<1>, dim3(grid), dim3(256), 0, s, d_pts, d_inf, d_scalars, (uint32_t)n, d_pts2,
                           d_inf2, d_scalars2);
    else if (curve == 2)
        hipLaunchKernelGGL(k_glv_expand<2>, dim3(grid), dim3(256), 0, s, d_pts, d_inf, d_scalars, (uint32_t)n, d_pts2,
                           d_inf2, d_scalars2);
    else
        throw std::invalid_argument("GLV split: unknown curve");
    KZ_HIP(hipGetLastError());
}

void check_g1(int curve, const uint32_t* d_pts, const uint8_t* d_inf, size_t n, int* d_bad, hipStream_t s) {
    if (n == 0) return;
    const unsigned grid = ceil_div(n, 64);
    if (curve == 2)
        hipLaunchKernelGGL(k_check_g1<CurveBLS12381>, dim3(grid), dim3(64), 0, s, d_pts, d_inf, (uint32_t)n, d_bad);
    else if (curve == 1)
        hipLaunchKernelGGL(k_check_g1<CurveBN254>, dim3(grid), dim3(64), 0, s, d_pts, d_inf, (uint32_t)n, d_bad);
    else
        hipLaunchKernelGGL(k_check_g1<CurveSecp256k1>, dim3(grid), dim3(64), 0, s, d_pts, d_inf, (uint32_t)n, d_bad);
    KZ_HIP(hipGetLastError());
}

}  // namespace kz

using namespace kz;

extern "C" {

int kinet_hip_points_check_g1(uint32_t curve, const void* d_pts, const void* d_inf, size_t n, int* d_bad,
                              void* stream) {
    if (curve > 2) return GPUKIT_ERR_BACKEND;
    if (!d_bad || (n > 0 && !d_pts)) return GPUKIT_ERR_NULL_ARG;
    try {
        (void)engine();
        check_g1((int)curve, static_cast<const uint32_t*>(d_pts), static_cast<const uint8_t*>(d_inf), n, d_bad,
                 static_cast<hipStream_t>(stream));
        return GPUKIT_OK;
    } catch (const std::exception& ex) {
        set_last_error(ex.what());
        return GPUKIT_ERR_BACKEND;
    }
}

}  // extern "C"
