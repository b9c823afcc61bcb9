// Quad-cooperative XYZZ point operations for latency-bound kernels.
//
// A lone serial chain of point operations (the window combine's 256
// doublings, a running sum) runs at one wave's instruction issue rate, so its
// time is (field products on the critical path) x (one product's latency).
// Here the four lanes of a quad (lanes 4k..4k+3) hold the same point(s) and
// evaluate one operation together: at each level of the formula's dependency
// graph lane q computes one field product with operands picked by q, and a
// quad-permute swizzle broadcasts the four products back across the quad.
// The lanes execute in lock step, so four products cost one product's
// latency:
//
//   doubling (dbl-2008-s-1)  9 products -> 3 levels
//   addition (add-2008-s)   14 products -> 4 levels
//
// The results are the field elements the serial formulas in curve.h
// produce: canonical ones under MulInline / MulCall, the same elements below
// 8p / 4p / 2p (X / Y / ZZ, ZZZ) under the lazy policies, so outputs stay
// bit-identical once made canonical (xyzz_canon_acc).
//
// Contract: all four lanes of a quad take the same branches (control flow
// depends only on the point data they share).  Different quads may diverge.
#pragma once

#include "curve.h"

namespace kz {

// Lane exchange inside a quad.  DeviceQuad reads lane K of each quad from
// every lane of that quad with ds_swizzle in quad-permute mode (offset bit 15
// set, pattern [K,K,K,K] in bits 7:0).  (A DPP quad_perm move does the same in
// a VALU slot, but the compiler folds such moves into the following
// subtraction, and that folded form gave wrong differences here; the swizzle
// is never folded.)  The point operations below take the exchange as a
// policy so that a host build can run them with four lock-stepped threads.
struct DeviceQuad {
    __device__ static int lane() { return (int)(threadIdx.x & 3u); }
    template <int K, class F>
    __device__ static Fe<F> read(const Fe<F>& a) {
        Fe<F> r;
#pragma unroll
        for (int i = 0; i < F::N; ++i) r.v[i] = (uint32_t)__builtin_amdgcn_ds_swizzle((int)a.v[i], 0x8000 | (K * 0x55));
        return r;
    }
};

__device__ __forceinline__ int quad_lane() { return DeviceQuad::lane(); }

template <int K, class F, class Lanes = DeviceQuad>
__host__ __device__ __forceinline__ Fe<F> quad_read(const Fe<F>& a) {
    return Lanes::template read<K, F>(a);
}

template <class F>
__host__ __device__ __forceinline__ Fe<F> quad_pick(int q, const Fe<F>& a0, const Fe<F>& a1, const Fe<F>& a2,
                                                    const Fe<F>& a3) {
    Fe<F> r;
#pragma unroll
    for (int i = 0; i < F::N; ++i) {
        const uint32_t lo = q == 0 ? a0.v[i] : a1.v[i];
        const uint32_t hi = q == 2 ? a2.v[i] : a3.v[i];
        r.v[i] = q < 2 ? lo : hi;
    }
    return r;
}

// 2P, three product levels.  (The curves' groups have odd order: doubling a
// non-identity point never gives the identity, so ZZ stays non-zero.)
//
// Under a lazy policy the coordinates keep the accumulation loop's bounds
// (field.h): X < 8p, Y < 4p, ZZ, ZZZ < 2p on entry and exit, differences in
// one pass (a + k p - b), every product operand pair at most 64 p^2; the
// caller makes stored values canonical with xyzz_canon_acc.
template <class F, class M = MulCall, class Lanes = DeviceQuad>
__host__ __device__ Xyzz<F> xyzz_dbl_quad(const Xyzz<F>& p) {
    constexpr bool L = M::kLazy;
    const int q = Lanes::lane();
    const Fe<F> U = L ? fe_small_mul<F, 2>(p.Y) : fe_dbl<F>(p.Y);  // lazy: < 8p
    // level 1: V = U^2, XX = X^2
    Fe<F> m = M::template mul<F>(quad_pick<F>(q, U, p.X, U, p.X), quad_pick<F>(q, U, p.X, U, p.X));
    const Fe<F> V = quad_read<0, F, Lanes>(m), XX = quad_read<1, F, Lanes>(m);
    const Fe<F> X3 = L ? fe_small_mul<F, 3>(XX) : fe_add<F>(fe_dbl<F>(XX), XX);  // 3 X^2 (lazy: < 6p)
    // level 2: W = U V, S = X V, ZZ' = V ZZ, (3 X^2)^2
    m = M::template mul<F>(quad_pick<F>(q, U, p.X, V, X3), quad_pick<F>(q, V, V, p.ZZ, X3));
    const Fe<F> W = quad_read<0, F, Lanes>(m), S = quad_read<1, F, Lanes>(m);
    Xyzz<F> r;
    r.ZZ = quad_read<2, F, Lanes>(m);
    r.X = L ? fe_kp_2b<F, 4>(quad_read<3, F, Lanes>(m), S) : fe_sub<F>(quad_read<3, F, Lanes>(m), fe_dbl<F>(S));  // lazy: < 6p
    // level 3: 3 X^2 (S - X'), W Y, W ZZZ
    const Fe<F> SmX = L ? fe_sub_kp<F, 6>(S, r.X) : fe_sub<F>(S, r.X);  // lazy: < 8p
    m = M::template mul<F>(quad_pick<F>(q, X3, W, W, W), quad_pick<F>(q, SmX, p.Y, p.ZZZ, p.ZZZ));
    r.Y = L ? fe_sub_kp<F, 2>(quad_read<0, F, Lanes>(m), quad_read<1, F, Lanes>(m)) : fe_sub<F>(quad_read<0, F, Lanes>(m), quad_read<1, F, Lanes>(m));
    r.ZZZ = quad_read<2, F, Lanes>(m);
    return r;
}

// a + b with the doubling and inverse special cases, four product levels.
// Lazy policy: the bounds of xyzz_dbl_quad; P, R < 4p, X' < 8p, Y' < 4p, and
// P or R is zero mod p exactly when PP or RR (products, below 2p) is 0 or p,
// so the special cases are decided after the second level.
template <class F, class M = MulCall, class Lanes = DeviceQuad>
__host__ __device__ Xyzz<F> xyzz_add_quad(const Xyzz<F>& a, const Xyzz<F>& b) {
    constexpr bool L = M::kLazy;
    if (xyzz_is_inf<F>(a)) return b;
    if (xyzz_is_inf<F>(b)) return a;
    const int q = Lanes::lane();
    // level 1: U1 = X1 ZZ2, U2 = X2 ZZ1, S1 = Y1 ZZZ2, S2 = Y2 ZZZ1
    Fe<F> m = M::template mul<F>(quad_pick<F>(q, a.X, b.X, a.Y, b.Y), quad_pick<F>(q, b.ZZ, a.ZZ, b.ZZZ, a.ZZZ));
    const Fe<F> U1 = quad_read<0, F, Lanes>(m), S1 = quad_read<2, F, Lanes>(m);
    const Fe<F> P = L ? fe_sub_kp<F, 2>(quad_read<1, F, Lanes>(m), U1) : fe_sub<F>(quad_read<1, F, Lanes>(m), U1);
    const Fe<F> R = L ? fe_sub_kp<F, 2>(quad_read<3, F, Lanes>(m), S1) : fe_sub<F>(quad_read<3, F, Lanes>(m), S1);
    if constexpr (!L) {
        if (fe_is_zero<F>(P)) {
            if (fe_is_zero<F>(R)) return xyzz_dbl_quad<F, M, Lanes>(a);
            return xyzz_inf<F>();
        }
    }
    // level 2: PP = P^2, RR = R^2, ZZ1 ZZ2, ZZZ1 ZZZ2
    m = M::template mul<F>(quad_pick<F>(q, P, R, a.ZZ, a.ZZZ), quad_pick<F>(q, P, R, b.ZZ, b.ZZZ));
    const Fe<F> PP = quad_read<0, F, Lanes>(m), RR = quad_read<1, F, Lanes>(m), Z12 = quad_read<2, F, Lanes>(m), ZZZ12 = quad_read<3, F, Lanes>(m);
    if constexpr (L) {
        if (fe_is_zero_lazy<F>(PP)) {
            if (fe_is_zero_lazy<F>(RR)) return xyzz_dbl_quad<F, M, Lanes>(a);
            return xyzz_inf<F>();
        }
    }
    // level 3: PPP = P PP, Q = U1 PP, ZZ' = ZZ1 ZZ2 PP
    m = M::template mul<F>(quad_pick<F>(q, P, U1, Z12, Z12), quad_pick<F>(q, PP, PP, PP, PP));
    const Fe<F> PPP = quad_read<0, F, Lanes>(m), Q = quad_read<1, F, Lanes>(m);
    Xyzz<F> r;
    r.ZZ = quad_read<2, F, Lanes>(m);
    r.X = L ? fe_r6p_s_2q<F>(RR, PPP, Q) : fe_sub<F>(fe_sub<F>(RR, PPP), fe_dbl<F>(Q));
    // level 4: R (Q - X'), S1 PPP, ZZZ' = ZZZ1 ZZZ2 PPP
    const Fe<F> QmX = L ? fe_sub_kp<F, 8>(Q, r.X) : fe_sub<F>(Q, r.X);
    m = M::template mul<F>(quad_pick<F>(q, R, S1, ZZZ12, ZZZ12), quad_pick<F>(q, QmX, PPP, PPP, PPP));
    r.Y = L ? fe_sub_kp<F, 2>(quad_read<0, F, Lanes>(m), quad_read<1, F, Lanes>(m)) : fe_sub<F>(quad_read<0, F, Lanes>(m), quad_read<1, F, Lanes>(m));
    r.ZZZ = quad_read<2, F, Lanes>(m);
    return r;
}

}  // namespace kz
