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
// produce: canonical ones under MulInline / MulCall, the same elements in
// [0, 2p) under the lazy policies, so outputs stay bit-identical once made
// canonical.
//
// Contract: all four lanes of a quad take the same branches (control flow
// depends only on the point data they share).  Different quads may diverge.
#pragma once

#include "curve.h"

namespace kz {

__device__ __forceinline__ int quad_lane() { return (int)(threadIdx.x & 3u); }

// Lane K of each quad, read by every lane of that quad: ds_swizzle in
// quad-permute mode (offset bit 15 set, pattern [K,K,K,K] in bits 7:0).
// (A DPP quad_perm move does the same in a VALU slot, but the compiler folds
// such moves into the following subtraction, and that folded form gave wrong
// differences here; the swizzle is never folded.)
template <int K>
__device__ __forceinline__ uint32_t quad_read_u32(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, 0x8000 | (K * 0x55));
}

template <int K, class F>
__device__ __forceinline__ Fe<F> quad_read(const Fe<F>& a) {
    Fe<F> r;
#pragma unroll
    for (int i = 0; i < F::N; ++i) r.v[i] = quad_read_u32<K>(a.v[i]);
    return r;
}

template <class F>
__device__ __forceinline__ Fe<F> quad_pick(int q, const Fe<F>& a0, const Fe<F>& a1, const Fe<F>& a2,
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

// Additive helpers matching the product policy: canonical values, or with a
// lazy policy (MulLazy, MulCallLazy) values in [0, 2p) (field.h), which the
// caller makes canonical (xyzz_canon) before storing or comparing them.
template <class F, bool L>
__device__ __forceinline__ Fe<F> q_sub(const Fe<F>& a, const Fe<F>& b) {
    if constexpr (L) return fe_sub_lazy<F>(a, b);
    else return fe_sub<F>(a, b);
}
template <class F, bool L>
__device__ __forceinline__ Fe<F> q_add(const Fe<F>& a, const Fe<F>& b) {
    if constexpr (L) return fe_add_lazy<F>(a, b);
    else return fe_add<F>(a, b);
}
template <class F, bool L>
__device__ __forceinline__ Fe<F> q_dbl(const Fe<F>& a) {
    if constexpr (L) return fe_dbl_lazy<F>(a);
    else return fe_dbl<F>(a);
}
template <class F, bool L>
__device__ __forceinline__ bool q_is_zero(const Fe<F>& a) {
    if constexpr (L) return fe_is_zero_lazy<F>(a);
    else return fe_is_zero<F>(a);
}

// 2P, three product levels.  (The curves' groups have odd order: doubling a
// non-identity point never gives the identity, so ZZ stays non-zero.)
template <class F, class M = MulCall>
__device__ Xyzz<F> xyzz_dbl_quad(const Xyzz<F>& p) {
    constexpr bool L = M::kLazy;
    const int q = quad_lane();
    const Fe<F> U = q_dbl<F, L>(p.Y);
    // level 1: V = U^2, XX = X^2
    Fe<F> m = M::template mul<F>(quad_pick<F>(q, U, p.X, U, p.X), quad_pick<F>(q, U, p.X, U, p.X));
    const Fe<F> V = quad_read<0>(m), XX = quad_read<1>(m);
    const Fe<F> X3 = q_add<F, L>(q_dbl<F, L>(XX), XX);  // 3 X^2
    // level 2: W = U V, S = X V, ZZ' = V ZZ, (3 X^2)^2
    m = M::template mul<F>(quad_pick<F>(q, U, p.X, V, X3), quad_pick<F>(q, V, V, p.ZZ, X3));
    const Fe<F> W = quad_read<0>(m), S = quad_read<1>(m);
    Xyzz<F> r;
    r.ZZ = quad_read<2>(m);
    r.X = q_sub<F, L>(quad_read<3>(m), q_dbl<F, L>(S));
    // level 3: 3 X^2 (S - X'), W Y, W ZZZ
    m = M::template mul<F>(quad_pick<F>(q, X3, W, W, W), quad_pick<F>(q, q_sub<F, L>(S, r.X), p.Y, p.ZZZ, p.ZZZ));
    r.Y = q_sub<F, L>(quad_read<0>(m), quad_read<1>(m));
    r.ZZZ = quad_read<2>(m);
    return r;
}

// a + b with the doubling and inverse special cases, four product levels.
template <class F, class M = MulCall>
__device__ Xyzz<F> xyzz_add_quad(const Xyzz<F>& a, const Xyzz<F>& b) {
    constexpr bool L = M::kLazy;
    if (xyzz_is_inf<F>(a)) return b;
    if (xyzz_is_inf<F>(b)) return a;
    const int q = quad_lane();
    // level 1: U1 = X1 ZZ2, U2 = X2 ZZ1, S1 = Y1 ZZZ2, S2 = Y2 ZZZ1
    Fe<F> m = M::template mul<F>(quad_pick<F>(q, a.X, b.X, a.Y, b.Y), quad_pick<F>(q, b.ZZ, a.ZZ, b.ZZZ, a.ZZZ));
    const Fe<F> U1 = quad_read<0>(m), S1 = quad_read<2>(m);
    const Fe<F> P = q_sub<F, L>(quad_read<1>(m), U1);
    const Fe<F> R = q_sub<F, L>(quad_read<3>(m), S1);
    if (q_is_zero<F, L>(P)) {
        if (q_is_zero<F, L>(R)) return xyzz_dbl_quad<F, M>(a);
        return xyzz_inf<F>();
    }
    // level 2: PP = P^2, RR = R^2, ZZ1 ZZ2, ZZZ1 ZZZ2
    m = M::template mul<F>(quad_pick<F>(q, P, R, a.ZZ, a.ZZZ), quad_pick<F>(q, P, R, b.ZZ, b.ZZZ));
    const Fe<F> PP = quad_read<0>(m), RR = quad_read<1>(m), Z12 = quad_read<2>(m), ZZZ12 = quad_read<3>(m);
    // level 3: PPP = P PP, Q = U1 PP, ZZ' = ZZ1 ZZ2 PP
    m = M::template mul<F>(quad_pick<F>(q, P, U1, Z12, Z12), quad_pick<F>(q, PP, PP, PP, PP));
    const Fe<F> PPP = quad_read<0>(m), Q = quad_read<1>(m);
    Xyzz<F> r;
    r.ZZ = quad_read<2>(m);
    r.X = q_sub<F, L>(q_sub<F, L>(RR, PPP), q_dbl<F, L>(Q));
    // level 4: R (Q - X'), S1 PPP, ZZZ' = ZZZ1 ZZZ2 PPP
    m = M::template mul<F>(quad_pick<F>(q, R, S1, ZZZ12, ZZZ12), quad_pick<F>(q, q_sub<F, L>(Q, r.X), PPP, PPP, PPP));
    r.Y = q_sub<F, L>(quad_read<0>(m), quad_read<1>(m));
    r.ZZZ = quad_read<2>(m);
    return r;
}

}  // namespace kz
