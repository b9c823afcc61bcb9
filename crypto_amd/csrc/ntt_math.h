// Butterflies of the scalar-field NTT with lazy reduction, host- and
// device-compilable (the host check runs the same functions over whole
// transforms on the CPU).
//
// Inside a pass no value is reduced mod r: a product skips the final
// conditional subtraction, a sum is a + b and a difference is a + 2r - b,
// limbs kept normalised (< 2^29) by one carry pass.  Bounds, for a pass whose
// inputs are below 2r (canonical inputs, or a previous pass's products):
//   - a Montgomery product a * w with w < r (a twiddle, canonical) is below
//     r + a w / R = r (1 + a / R);
//   - a twiddle-free stage (the first) takes a bound B to 2B = 4r, every
//     stage with twiddles to B + 2r, since its products stay below 2r;
//   - a pass of 2^11 points therefore ends below 24r (2^10: 22r).
// With R = 2^261, R / r is 70.7 for BLS12-381 Fr and 169 for BN254 Fr, so
// every operand stays far below R (products below 1.4r) and every value fits
// nine normalised limbs.  Differences a + 2r - b need b < 2r: the
// twiddle-free stage subtracts a pass input (< 2r), every other stage a
// product.  A pass ends with a product (the four-step twiddle, or n^-1: below
// 2r) or with fe_reduce_wide (canonical).
#pragma once

#include "field.h"

namespace kz {

// x mod p for a normalised x below 64p.  q = floor(x_top / (p_top + 1)) from
// the top limbs never exceeds floor(x / p) and falls short of it by at most
// one (p_top >= 2^21 for the 9-limb fields, so (x_top + p_top + 1) /
// (p_top (p_top + 1)) < 1), so x - q p lies in [0, 2p) and one conditional
// subtraction finishes.
template <class F>
KZ_DEV Fe<F> fe_reduce_wide(const Fe<F>& x) {
    constexpr uint32_t top = F::P[F::N - 1] + 1u;
    const uint32_t q = x.v[F::N - 1] / top;
    Fe<F> d;
    int64_t c = 0;
#pragma unroll
    for (int i = 0; i < F::N; ++i) {
        const int64_t t = (int64_t)x.v[i] - (int64_t)q * F::P[i] + c;
        c = t >> kLimbBits;
        d.v[i] = (uint32_t)t & kLimbMask;
    }
    fe_reduce_once<F>(d);
    return d;
}

// Montgomery product without the final subtraction: below p (1 + a / R) for
// a canonical b.
template <class F>
KZ_DEV Fe<F> fe_mul_lazy(const Fe<F>& a, const Fe<F>& b) {
    return fe_mul_k<F, 1, false>(a, b);
}

// Twiddle-free radix-2 butterfly: (u, v) -> (u + v, u + 2p - v), v < 2p.
template <class F>
KZ_DEV void ntt_bfly0(Fe<F>& u, Fe<F>& v) {
    const Fe<F> a = fe_add_nored<F>(u, v);
    v = fe_sub_kp<F, 2>(u, v);
    u = a;
}

// Radix-2 stages s and s + 1 fused over the four elements e, e + h, e + 2h,
// e + 3h (h = 2^s) that they combine among themselves.  Stage s multiplies
// x1 and x3 by ws = w_(2h)^i (Stage0: stage s is the twiddle-free first
// stage); stage s + 1 pairs (e, e + 2h) with w1 = w_(4h)^i and (e + h, e + 3h)
// with w2 = w_(4h)^(i + h), multiplying after the sum and difference.
template <class F, bool Stage0>
KZ_DEV void ntt_radix4(Fe<F>& x0, Fe<F>& x1, Fe<F>& x2, Fe<F>& x3, const Fe<F>& ws, const Fe<F>& w1,
                       const Fe<F>& w2) {
    if (!Stage0) {
        x1 = fe_mul_lazy<F>(x1, ws);
        x3 = fe_mul_lazy<F>(x3, ws);
    }
    const Fe<F> a0 = fe_add_nored<F>(x0, x1), a1 = fe_sub_kp<F, 2>(x0, x1);
    const Fe<F> t0 = fe_mul_lazy<F>(fe_add_nored<F>(x2, x3), w1);
    const Fe<F> t1 = fe_mul_lazy<F>(fe_sub_kp<F, 2>(x2, x3), w2);
    x0 = fe_add_nored<F>(a0, t0);
    x2 = fe_sub_kp<F, 2>(a0, t0);
    x1 = fe_add_nored<F>(a1, t1);
    x3 = fe_sub_kp<F, 2>(a1, t1);
}

}  // namespace kz
