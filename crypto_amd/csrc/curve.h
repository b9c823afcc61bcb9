// Device G1 arithmetic for short-Weierstrass curves with a = 0 (BLS12-381 G1,
// BN254 G1), in XYZZ coordinates: x = X/ZZ, y = Y/ZZZ with ZZ^3 = ZZZ^2.
//
// XYZZ ("extended Jacobian") needs no inversion, its mixed addition is 8M+2S
// against 7M+4S for Jacobian, and infinity is simply ZZ = ZZZ = 0.  The group
// law (and hence every MSM result) is the same as the reference's Jacobian
// add-2007-bl / dbl-2009-l (bls/cpp/bls12_381_g1.hpp:74-151); only the
// projective representation differs, and outputs are normalised to affine.
//
// Formulas: Bernstein-Lange EFD g1p/short-w "xyzz": madd-2008-s, add-2008-s,
// dbl-2008-s-1, mdbl-2008-s-1.
#pragma once

#include "field.h"

namespace kz {

struct CurveBN254 {
    using Fp = BnFp;
    using Fr = BnFr;
    static constexpr int FB = 32;        // bytes per coordinate on the wire
    static constexpr uint32_t B = 3;
    static constexpr int ID = 1;         // gpukit_curve
};

struct CurveBLS12381 {
    using Fp = BlsFp;
    using Fr = BlsFr;
    static constexpr int FB = 48;
    static constexpr uint32_t B = 4;
    static constexpr int ID = 2;
};

// secp256k1 (y^2 = x^3 + 7, secp256k1/cpp/curve.hpp): only the plain MSM of
// the multi-curve entry runs on it, so it has no scalar-field type here.
struct CurveSecp256k1 {
    using Fp = SecpFp;
    using Fr = void;
    static constexpr int FB = 32;
    static constexpr uint32_t B = 7;
    static constexpr int ID = 0;
};

template <class F>
struct Affine {
    Fe<F> x, y;
};

template <class F>
struct Xyzz {
    Fe<F> X, Y, ZZ, ZZZ;
};

// Multiplication policy of the point operations.  MulInline expands every
// field product in place (the hot accumulation loop); MulCall calls one
// out-of-line copy, keeping latency-bound kernels (bucket reduction, window
// combine) inside the instruction cache.
//
// The out-of-line copy takes its operands as 2 x 14 plain 32-bit arguments
// (limbs past N are zero), which the calling convention keeps in VGPRs, and
// returns the product in VGPRs: no stack traffic.  (Passed as two Fe structs,
// the second one would go through the stack as a byval copy in scratch.)
template <class F, bool Canon = true>
__host__ __device__ __noinline__ Fe<F> fe_mul_flat(uint32_t a0, uint32_t a1, uint32_t a2, uint32_t a3, uint32_t a4,
                                                  uint32_t a5, uint32_t a6, uint32_t a7, uint32_t a8, uint32_t a9,
                                                  uint32_t a10, uint32_t a11, uint32_t a12, uint32_t a13,
                                                  uint32_t b0, uint32_t b1, uint32_t b2, uint32_t b3, uint32_t b4,
                                                  uint32_t b5, uint32_t b6, uint32_t b7, uint32_t b8, uint32_t b9,
                                                  uint32_t b10, uint32_t b11, uint32_t b12, uint32_t b13) {
    const uint32_t av[14] = {a0, a1, a2, a3, a4, a5, a6, a7, a8, a9, a10, a11, a12, a13};
    const uint32_t bv[14] = {b0, b1, b2, b3, b4, b5, b6, b7, b8, b9, b10, b11, b12, b13};
    Fe<F> a, b;
    for (int i = 0; i < F::N; ++i) {
        a.v[i] = av[i];
        b.v[i] = bv[i];
    }
    return fe_mul_k<F, KZ_MUL_ILP, Canon>(a, b);
}

template <int I, class F>
KZ_DEV uint32_t limb_or_zero(const Fe<F>& a) {
    if constexpr (I < F::N) return a.v[I];
    else return 0u;
}

template <class F, bool Canon = true>
KZ_DEV Fe<F> fe_mul_call(const Fe<F>& a, const Fe<F>& b) {
    return fe_mul_flat<F, Canon>(limb_or_zero<0>(a), limb_or_zero<1>(a), limb_or_zero<2>(a), limb_or_zero<3>(a),
                          limb_or_zero<4>(a), limb_or_zero<5>(a), limb_or_zero<6>(a), limb_or_zero<7>(a),
                          limb_or_zero<8>(a), limb_or_zero<9>(a), limb_or_zero<10>(a), limb_or_zero<11>(a),
                          limb_or_zero<12>(a), limb_or_zero<13>(a), limb_or_zero<0>(b), limb_or_zero<1>(b),
                          limb_or_zero<2>(b), limb_or_zero<3>(b), limb_or_zero<4>(b), limb_or_zero<5>(b),
                          limb_or_zero<6>(b), limb_or_zero<7>(b), limb_or_zero<8>(b), limb_or_zero<9>(b),
                          limb_or_zero<10>(b), limb_or_zero<11>(b), limb_or_zero<12>(b), limb_or_zero<13>(b));
}

struct MulInline {
    static constexpr bool kLazy = false;
    template <class F>
    static KZ_DEV Fe<F> mul(const Fe<F>& a, const Fe<F>& b) {
        return fe_mul<F>(a, b);
    }
    template <class F>
    static KZ_DEV Fe<F> sqr(const Fe<F>& a) {
        return fe_sqr<F>(a);
    }
};

struct MulCall {
    static constexpr bool kLazy = false;
    template <class F>
    static KZ_DEV Fe<F> mul(const Fe<F>& a, const Fe<F>& b) {
        return fe_mul_call<F>(a, b);
    }
    template <class F>
    static KZ_DEV Fe<F> sqr(const Fe<F>& a) {  // the one out-of-line product
        return fe_mul_call<F>(a, a);
    }
};

// Products without the final subtraction (the lazy-reduction section of
// field.h): results below 2p for operand pairs below R p.  Only for chains
// that keep the bounds of xyzz_madd_lazy / the lazy quad operations.
struct MulLazy {
    static constexpr bool kLazy = true;
    template <class F>
    static KZ_DEV Fe<F> mul(const Fe<F>& a, const Fe<F>& b) {
        return fe_mul_k<F, KZ_MUL_ILP, false>(a, b);
    }
    template <class F>
    static KZ_DEV Fe<F> sqr(const Fe<F>& a) {
        return fe_sqr<F, false>(a);
    }
};

// The out-of-line product without its final subtraction.
struct MulCallLazy {
    static constexpr bool kLazy = true;
    template <class F>
    static KZ_DEV Fe<F> mul(const Fe<F>& a, const Fe<F>& b) {
        return fe_mul_call<F, false>(a, b);
    }
    template <class F>
    static KZ_DEV Fe<F> sqr(const Fe<F>& a) {
        return fe_mul_call<F, false>(a, a);
    }
};

template <class F>
KZ_DEV Xyzz<F> xyzz_inf() {
    Xyzz<F> r;
    r.X = fe_zero<F>();
    r.Y = fe_zero<F>();
    r.ZZ = fe_zero<F>();
    r.ZZZ = fe_zero<F>();
    return r;
}

template <class F>
KZ_DEV bool xyzz_is_inf(const Xyzz<F>& p) {
    return fe_is_zero<F>(p.ZZ);
}

template <class F>
KZ_DEV Xyzz<F> xyzz_from_affine(const Affine<F>& a) {
    Xyzz<F> r;
    r.X = a.x;
    r.Y = a.y;
    r.ZZ = fe_one<F>();
    r.ZZZ = fe_one<F>();
    return r;
}

template <class F>
KZ_DEV Xyzz<F> xyzz_neg(const Xyzz<F>& p) {
    Xyzz<F> r = p;
    r.Y = fe_neg<F>(p.Y);
    return r;
}

// 2*(x, y) for an affine point (mdbl-2008-s-1).
template <class F, class M = MulInline>
KZ_DEV Xyzz<F> xyzz_dbl_affine(const Affine<F>& a) {
    Fe<F> U = fe_dbl<F>(a.y);
    Fe<F> V = M::template mul<F>(U, U);
    Fe<F> W = M::template mul<F>(U, V);
    Fe<F> S = M::template mul<F>(a.x, V);
    Fe<F> xx = M::template mul<F>(a.x, a.x);
    Fe<F> xx3 = fe_add<F>(fe_dbl<F>(xx), xx);
    Xyzz<F> r;
    r.X = fe_sub<F>(M::template mul<F>(xx3, xx3), fe_dbl<F>(S));
    r.Y = fe_sub<F>(M::template mul<F>(xx3, fe_sub<F>(S, r.X)), M::template mul<F>(W, a.y));
    r.ZZ = V;
    r.ZZZ = W;
    return r;
}

// 2*P (dbl-2008-s-1, a = 0).  Y = 0 gives V = 0, i.e. infinity.
template <class F, class M = MulInline>
KZ_DEV Xyzz<F> xyzz_dbl(const Xyzz<F>& p) {
    Fe<F> U = fe_dbl<F>(p.Y);
    Fe<F> V = M::template sqr<F>(U);
    Fe<F> W = M::template mul<F>(U, V);
    Fe<F> S = M::template mul<F>(p.X, V);
    Fe<F> xx = M::template sqr<F>(p.X);
    Fe<F> xx3 = fe_add<F>(fe_dbl<F>(xx), xx);
    Xyzz<F> r;
    r.X = fe_sub<F>(M::template sqr<F>(xx3), fe_dbl<F>(S));
    r.Y = fe_sub<F>(M::template mul<F>(xx3, fe_sub<F>(S, r.X)), M::template mul<F>(W, p.Y));
    r.ZZ = M::template mul<F>(V, p.ZZ);
    r.ZZZ = M::template mul<F>(W, p.ZZZ);
    return r;
}

// acc + (x, y), (x, y) an affine point that is not infinity (madd-2008-s).
// The rare equal-points branch stays inline: an out-of-line call there makes
// the caller save registers around it and spill in the hot loop.
template <class F, class M = MulInline>
KZ_DEV void xyzz_madd(Xyzz<F>& acc, const Affine<F>& q) {
    if (xyzz_is_inf<F>(acc)) {
        acc = xyzz_from_affine<F>(q);
        return;
    }
    Fe<F> U2 = M::template mul<F>(q.x, acc.ZZ);
    Fe<F> S2 = M::template mul<F>(q.y, acc.ZZZ);
    Fe<F> P = fe_sub<F>(U2, acc.X);
    Fe<F> R = fe_sub<F>(S2, acc.Y);
    if (fe_is_zero<F>(P)) {
        if (fe_is_zero<F>(R)) acc = xyzz_dbl_affine<F, M>(q);
        else acc = xyzz_inf<F>();
        return;
    }
    Fe<F> PP = M::template sqr<F>(P);
    Fe<F> PPP = M::template mul<F>(P, PP);
    Fe<F> Q = M::template mul<F>(acc.X, PP);
    Fe<F> X3 = fe_sub<F>(fe_sub<F>(M::template sqr<F>(R), PPP), fe_dbl<F>(Q));
    Fe<F> Y3 = fe_sub<F>(M::template mul<F>(R, fe_sub<F>(Q, X3)), M::template mul<F>(acc.Y, PPP));
    acc.X = X3;
    acc.Y = Y3;
    acc.ZZ = M::template mul<F>(acc.ZZ, PP);
    acc.ZZZ = M::template mul<F>(acc.ZZZ, PPP);
}

// xyzz_madd for the accumulation loop, with wider bounds (field.h): on entry
// acc.X < 8p, acc.Y < 4p, acc.ZZ, acc.ZZZ < 2p (canonical values included),
// q canonical; the result keeps the same bounds, and xyzz_canon_acc makes it
// canonical.  The same point as xyzz_madd:
//   P = U2 + 8p - X (< 10p), R = S2 + 4p - Y (< 6p), PP = P^2, RR = R^2,
//   X3 = RR + 6p - PPP - 2Q (< 8p), Y3 = R (Q + 8p - X3) + 2p - Y PPP (< 4p),
// every product operand pair below 100 p^2.  P or R is zero mod p exactly
// when PP or RR (products, below 2p) is 0 or p.  The identity is only ever the
// canonical xyzz_inf (ZZ is a product of non-zero values otherwise), so
// xyzz_is_inf applies unchanged.  In the doubling case acc equals q and is
// doubled canonically: q is dead after its two products, which keeps the
// accumulation loop within 168 registers without spills.
template <class F>
KZ_DEV Xyzz<F> xyzz_canon_acc(Xyzz<F> p) {
    fe_csub_kp<F, 4>(p.X);
    fe_csub_kp<F, 2>(p.X);
    fe_csub_kp<F, 1>(p.X);
    fe_csub_kp<F, 2>(p.Y);
    fe_csub_kp<F, 1>(p.Y);
    fe_reduce_once<F>(p.ZZ);
    fe_reduce_once<F>(p.ZZZ);
    return p;
}

template <class F>
KZ_DEV void xyzz_madd_lazy(Xyzz<F>& acc, const Affine<F>& q) {
    using M = MulLazy;
    if (xyzz_is_inf<F>(acc)) {
        acc = xyzz_from_affine<F>(q);
        return;
    }
    const Fe<F> P = fe_sub_kp<F, 8>(M::template mul<F>(q.x, acc.ZZ), acc.X);
    const Fe<F> R = fe_sub_kp<F, 4>(M::template mul<F>(q.y, acc.ZZZ), acc.Y);
    const Fe<F> PP = M::template sqr<F>(P);
    if (fe_is_zero_lazy<F>(PP)) {
        if (fe_is_zero_lazy<F>(M::template sqr<F>(R))) acc = xyzz_dbl<F, MulInline>(xyzz_canon_acc<F>(acc));  // 2q
        else acc = xyzz_inf<F>();
        return;
    }
    const Fe<F> PPP = M::template mul<F>(P, PP);
    acc.ZZZ = M::template mul<F>(acc.ZZZ, PPP);
    const Fe<F> Q = M::template mul<F>(acc.X, PP);
    acc.ZZ = M::template mul<F>(acc.ZZ, PP);
    const Fe<F> X3 = fe_r6p_s_2q<F>(M::template sqr<F>(R), PPP, Q);
    const Fe<F> nPPP = fe_sub_kp<F, 2>(fe_zero<F>(), PPP);
    // Y3 = R (Q - X3) - Y1 PPP as R (Q + 8p - X3) + Y1 (2p - PPP): one
    // Montgomery reduction for both products (R < 6p, Q + 8p - X3 < 10p,
    // Y1 < 4p, 2p - PPP <= 2p: the sum is far below p 2^(29N), so Y3 < 2p)
    acc.Y = fe_mul2_lazy<F>(R, fe_sub_kp<F, 8>(Q, X3), acc.Y, nPPP);
    acc.X = X3;
}

// a + b, general (add-2008-s), with the doubling and inverse special cases.
template <class F, class M = MulInline>
KZ_DEV Xyzz<F> xyzz_add(const Xyzz<F>& a, const Xyzz<F>& b) {
    if (xyzz_is_inf<F>(a)) return b;
    if (xyzz_is_inf<F>(b)) return a;
    Fe<F> U1 = M::template mul<F>(a.X, b.ZZ);
    Fe<F> U2 = M::template mul<F>(b.X, a.ZZ);
    Fe<F> S1 = M::template mul<F>(a.Y, b.ZZZ);
    Fe<F> S2 = M::template mul<F>(b.Y, a.ZZZ);
    Fe<F> P = fe_sub<F>(U2, U1);
    Fe<F> R = fe_sub<F>(S2, S1);
    if (fe_is_zero<F>(P)) {
        if (fe_is_zero<F>(R)) return xyzz_dbl<F, M>(a);
        return xyzz_inf<F>();
    }
    Fe<F> PP = M::template sqr<F>(P);
    Fe<F> PPP = M::template mul<F>(P, PP);
    Fe<F> Q = M::template mul<F>(U1, PP);
    Xyzz<F> r;
    r.X = fe_sub<F>(fe_sub<F>(M::template sqr<F>(R), PPP), fe_dbl<F>(Q));
    r.Y = fe_sub<F>(M::template mul<F>(R, fe_sub<F>(Q, r.X)), M::template mul<F>(S1, PPP));
    r.ZZ = M::template mul<F>(M::template mul<F>(a.ZZ, b.ZZ), PP);
    r.ZZZ = M::template mul<F>(M::template mul<F>(a.ZZZ, b.ZZZ), PPP);
    return r;
}

// Affine normalisation (x = X/ZZ, y = Y/ZZZ) with one inversion of ZZ*ZZZ.
template <class F>
KZ_DEV void xyzz_to_affine(const Xyzz<F>& p, Affine<F>& out, bool& inf) {
    inf = xyzz_is_inf<F>(p);
    if (inf) {
        out.x = fe_zero<F>();
        out.y = fe_zero<F>();
        return;
    }
    Fe<F> I = fe_inv<F>(fe_mul_call<F>(p.ZZ, p.ZZZ));   // Z^-5
    Fe<F> zz_inv = fe_mul_call<F>(I, p.ZZZ);             // Z^-2
    Fe<F> zzz_inv = fe_mul_call<F>(I, p.ZZ);             // Z^-3
    out.x = fe_mul_call<F>(p.X, zz_inv);
    out.y = fe_mul_call<F>(p.Y, zzz_inv);
}

// ------------------------------------------------------------ memory I/O
// Point records in HBM are padded to a multiple of 4 words so every record is
// 16-byte aligned and moves as whole dwordx4 transfers:
//   affine  x[N] y[N]            -> kAffWords  (BLS12-381 28, BN254 20) of
//                                   data in a kAffStride slot (32, 20)
//   XYZZ    X[N] Y[N] ZZ[N] ZZZ[N] -> kXyzzWords (BLS12-381 56, BN254 36)
// Affine records sit kAffStride words apart: BLS12-381's 112-byte records
// take a 128-byte slot, so a random gather touches one cache line instead of
// two about half the time (the accumulation gathers every point once per window).
template <class F>
struct Rec {
    static constexpr int kAffWords = (2 * F::N + 3) / 4 * 4;
    static constexpr int kAffStride = kAffWords == 28 ? 32 : kAffWords;
    static constexpr int kXyzzWords = (4 * F::N + 3) / 4 * 4;
};

template <int K>
KZ_DEV void load_words(const uint32_t* __restrict__ p, uint32_t (&w)[K]) {
    static_assert(K % 4 == 0, "records are whole dwordx4");
    const uint4* q = reinterpret_cast<const uint4*>(p);
#pragma unroll
    for (int i = 0; i < K / 4; ++i) {
        const uint4 t = q[i];
        w[4 * i] = t.x;
        w[4 * i + 1] = t.y;
        w[4 * i + 2] = t.z;
        w[4 * i + 3] = t.w;
    }
}

template <int K>
KZ_DEV void store_words(uint32_t* __restrict__ p, const uint32_t (&w)[K]) {
    static_assert(K % 4 == 0, "records are whole dwordx4");
    uint4* q = reinterpret_cast<uint4*>(p);
#pragma unroll
    for (int i = 0; i < K / 4; ++i) q[i] = make_uint4(w[4 * i], w[4 * i + 1], w[4 * i + 2], w[4 * i + 3]);
}

template <class F>
KZ_DEV Affine<F> affine_load(const uint32_t* __restrict__ base, uint32_t idx) {
    constexpr int K = Rec<F>::kAffWords;
    uint32_t w[K];
    load_words<K>(base + (size_t)idx * Rec<F>::kAffStride, w);
    Affine<F> a;
#pragma unroll
    for (int i = 0; i < F::N; ++i) {
        a.x.v[i] = w[i];
        a.y.v[i] = w[F::N + i];
    }
    return a;
}

template <class F>
KZ_DEV void affine_store(uint32_t* __restrict__ base, size_t idx, const Affine<F>& a) {
    constexpr int K = Rec<F>::kAffStride;  // padding written as zeros
    uint32_t w[K];
#pragma unroll
    for (int i = 0; i < K; ++i) w[i] = 0;
#pragma unroll
    for (int i = 0; i < F::N; ++i) {
        w[i] = a.x.v[i];
        w[F::N + i] = a.y.v[i];
    }
    store_words<K>(base + idx * K, w);
}

template <class F>
KZ_DEV Xyzz<F> xyzz_load(const uint32_t* __restrict__ base, size_t idx) {
    constexpr int K = Rec<F>::kXyzzWords;
    uint32_t w[K];
    load_words<K>(base + idx * K, w);
    Xyzz<F> r;
#pragma unroll
    for (int i = 0; i < F::N; ++i) {
        r.X.v[i] = w[i];
        r.Y.v[i] = w[F::N + i];
        r.ZZ.v[i] = w[2 * F::N + i];
        r.ZZZ.v[i] = w[3 * F::N + i];
    }
    return r;
}

template <class F>
KZ_DEV void xyzz_store(uint32_t* __restrict__ base, size_t idx, const Xyzz<F>& a) {
    constexpr int K = Rec<F>::kXyzzWords;
    uint32_t w[K];
#pragma unroll
    for (int i = 0; i < K; ++i) w[i] = 0;
#pragma unroll
    for (int i = 0; i < F::N; ++i) {
        w[i] = a.X.v[i];
        w[F::N + i] = a.Y.v[i];
        w[2 * F::N + i] = a.ZZ.v[i];
        w[3 * F::N + i] = a.ZZZ.v[i];
    }
    store_words<K>(base + idx * K, w);
}

}  // namespace kz
