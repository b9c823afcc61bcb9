// Twisted Edwards arithmetic for Banderwagon (curve 3 of the multi-curve MSM):
// the Bandersnatch curve  a x^2 + y^2 = 1 + d x^2 y^2  over q = BLS12-381 r,
// a = -5, d = 0x6389c126...188d58e7 (banderwagon/cpp/element.hpp:6-22,
// element.cpp:22-45).  Host- and device-compilable, like curve.h.
//
// Extended coordinates (X : Y : T : Z), x = X/Z, y = Y/Z, T = XY/Z (Hisil,
// Wong, Carter, Dawson 2008), with the unified formulas of the EFD's
// twisted/extended set for a general a:
//   add-2008-hwcd   9M + 1 d-product;  mixed (Z2 = 1, T2 stored as d x2 y2) 8M
//   dbl-2008-hwcd   4S + 4M
// The reference sums with projective add-2008-bbjlp / dbl-2008-bbjlp
// (element.cpp:145-209).  Both are exact group laws on the curve: neither
// formula set has an exceptional case for the points the reader accepts (a
// point and its sums stay in the subgroup of order 2 r' that the Legendre test
// admits, where 1 +- d x1 x2 y1 y2 never vanishes), so every schedule gives
// the exact sum of the given representatives -- the (x, y) the reference's
// serialize_uncompressed writes, two-torsion component included.  The
// identity is (0 : 1 : 0 : 1); no operation needs a special case for it.
//
// Every value here is canonical (below q): the products are the full
// Montgomery products of field.h and the sums reduce; R / q is about 70.7 for
// the nine-limb representation, too little headroom for the wide lazy bounds
// the Weierstrass curves use.
#pragma once

#include "curve_quad.h"

namespace kz {

// Bandersnatch base field q = BLS12-381 r: BlsFr's constants under a type of
// its own, so that the MSM kernels select the Edwards point operations for it.
struct BwFp : BlsFr {};

struct CurveBanderwagon {
    using Fp = BwFp;
    using Fr = void;
    static constexpr int FB = 32;  // bytes per coordinate on the wire
    static constexpr int ID = 3;   // gpukit_curve
};

struct EdConst {
    // d R mod q (R = 2^261), radix-2^29 limbs
    static constexpr uint32_t D[9] = {0x1458E5F2u, 0x1CED1BB7u, 0x0C2440C6u, 0x03A6574Fu, 0x06EBC6F2u,
                                      0x05D944C8u, 0x185ECB02u, 0x1CDB6C09u, 0x006ED285u};
};

// a^((q-1)/2) == 1 (Euler's criterion: a is a non-zero square), by
// left-to-right square-and-multiply with the out-of-line product.
template <class F>
KZ_DEV bool fe_is_square_nonzero(const Fe<F>& a) {
    // (q - 1) / 2, little-endian 32-bit words
    constexpr uint32_t kHalf[8] = {0x80000000u, 0x7FFFFFFFu, 0x7FFF2DFFu, 0xA9DED201u,
                                   0x04D0EC02u, 0x199CEC04u, 0x94CEBEA4u, 0x39F6D3A9u};
    if (fe_is_zero<F>(a)) return false;
    Fe<F> r = a;  // the exponent's top bit is bit 253
#pragma unroll 1
    for (int bit = 252; bit >= 0; --bit) {
        r = fe_mul_call<F>(r, r);
        if ((kHalf[bit >> 5] >> (bit & 31)) & 1u) r = fe_mul_call<F>(r, a);
    }
    return fe_eq<F>(r, fe_one<F>());
}

template <class F>
struct EdExt {
    Fe<F> X, Y, T, Z;
};

// Affine point as the accumulation reads it: t = d x y.
template <class F>
struct EdAff {
    Fe<F> x, y, t;
};

template <class F>
KZ_DEV Fe<F> fe_mul5(const Fe<F>& a) {  // 5 a
    return fe_add<F>(fe_dbl<F>(fe_dbl<F>(a)), a);
}

template <class F>
KZ_DEV EdExt<F> ed_identity() {
    EdExt<F> r;
    r.X = fe_zero<F>();
    r.Y = fe_one<F>();
    r.T = fe_zero<F>();
    r.Z = fe_one<F>();
    return r;
}

template <class F>
KZ_DEV bool ed_is_identity(const EdExt<F>& p) {
    return fe_is_zero<F>(p.X) && fe_eq<F>(p.Y, p.Z);
}

template <class F, class M = MulInline>
KZ_DEV EdExt<F> ed_from_aff(const EdAff<F>& a) {
    EdExt<F> r;
    r.X = a.x;
    r.Y = a.y;
    r.T = M::template mul<F>(a.x, a.y);
    r.Z = fe_one<F>();
    return r;
}

// -(x, y) = (-x, y); t = d x y changes sign with x.
template <class F>
KZ_DEV EdAff<F> ed_neg_aff(EdAff<F> a) {
    a.x = fe_neg<F>(a.x);
    a.t = fe_neg<F>(a.t);
    return a;
}

template <class F>
KZ_DEV EdExt<F> ed_neg(EdExt<F> p) {
    p.X = fe_neg<F>(p.X);
    p.T = fe_neg<F>(p.T);
    return p;
}

// p + q, q affine (mixed add-2008-hwcd, Z2 = 1, q.t = d T2):
//   A = X1 x2, B = Y1 y2, C = T1 t2, E = (X1 + Y1)(x2 + y2) - A - B,
//   F = Z1 - C, G = Z1 + C, H = B - a A = B + 5 A,
//   X3 = E F, Y3 = G H, T3 = E H, Z3 = F G.
template <class F, class M = MulInline>
KZ_DEV void ed_madd(EdExt<F>& p, const EdAff<F>& q) {
    const Fe<F> A = M::template mul<F>(p.X, q.x);
    const Fe<F> B = M::template mul<F>(p.Y, q.y);
    const Fe<F> C = M::template mul<F>(p.T, q.t);
    const Fe<F> E = fe_sub<F>(fe_sub<F>(M::template mul<F>(fe_add<F>(p.X, p.Y), fe_add<F>(q.x, q.y)), A), B);
    const Fe<F> Fv = fe_sub<F>(p.Z, C), G = fe_add<F>(p.Z, C), H = fe_add<F>(B, fe_mul5<F>(A));
    p.X = M::template mul<F>(E, Fv);
    p.Y = M::template mul<F>(G, H);
    p.T = M::template mul<F>(E, H);
    p.Z = M::template mul<F>(Fv, G);
}

// ed_madd for the accumulation loop, lazily reduced (field.h's lazy section):
// on entry the accumulator's coordinates are below 2q (canonical ones
// included), the record canonical; the result keeps that bound and ed_canon
// makes it canonical.  A and B are full products (below q), so
//   X1 + Y1 < 4q, x2 + y2 < 2q, E = E' + 2q - A - B < 4q, F = Z1 + 2q - C < 4q,
//   G = Z1 + C < 4q, H = B + 5A < 6q,
// every product operand pair is at most 24 q^2 < R q (R / q = 70.7), and a
// lazy product of such a pair is below q (1 + 24 / 70.7) < 2q.
template <class F>
KZ_DEV void ed_madd_lazy(EdExt<F>& p, const EdAff<F>& q) {
    using L = MulLazy;
    const Fe<F> A = fe_mul<F>(p.X, q.x);
    const Fe<F> B = fe_mul<F>(p.Y, q.y);
    const Fe<F> C = L::template mul<F>(p.T, q.t);
    const Fe<F> Ep = L::template mul<F>(fe_add_nored<F>(p.X, p.Y), fe_add_nored<F>(q.x, q.y));
    const Fe<F> E = fe_kp_sub2<F, 2>(Ep, A, B);
    const Fe<F> Fv = fe_sub_kp<F, 2>(p.Z, C), G = fe_add_nored<F>(p.Z, C);
    const Fe<F> H = fe_add_nored<F>(B, fe_small_mul<F, 5>(A));
    p.X = L::template mul<F>(E, Fv);
    p.Y = L::template mul<F>(G, H);
    p.T = L::template mul<F>(E, H);
    p.Z = L::template mul<F>(Fv, G);
}

// Canonical coordinates of a point whose coordinates are below 2q.
template <class F>
KZ_DEV EdExt<F> ed_canon(EdExt<F> p) {
    fe_reduce_once<F>(p.X);
    fe_reduce_once<F>(p.Y);
    fe_reduce_once<F>(p.T);
    fe_reduce_once<F>(p.Z);
    return p;
}

// p + q (add-2008-hwcd): as ed_madd with C = d T1 T2 and D = Z1 Z2 in place of Z1.
template <class F, class M = MulInline>
KZ_DEV EdExt<F> ed_add(const EdExt<F>& p, const EdExt<F>& q) {
    const Fe<F> A = M::template mul<F>(p.X, q.X);
    const Fe<F> B = M::template mul<F>(p.Y, q.Y);
    const Fe<F> C = M::template mul<F>(M::template mul<F>(p.T, q.T), fe_const<F>(EdConst::D));
    const Fe<F> D = M::template mul<F>(p.Z, q.Z);
    const Fe<F> E = fe_sub<F>(fe_sub<F>(M::template mul<F>(fe_add<F>(p.X, p.Y), fe_add<F>(q.X, q.Y)), A), B);
    const Fe<F> Fv = fe_sub<F>(D, C), G = fe_add<F>(D, C), H = fe_add<F>(B, fe_mul5<F>(A));
    EdExt<F> r;
    r.X = M::template mul<F>(E, Fv);
    r.Y = M::template mul<F>(G, H);
    r.T = M::template mul<F>(E, H);
    r.Z = M::template mul<F>(Fv, G);
    return r;
}

// 2 p (dbl-2008-hwcd): A = X^2, B = Y^2, C = 2 Z^2, D = a A = -5 A,
// E = (X + Y)^2 - A - B, G = D + B, F = G - C, H = D - B,
// X3 = E F, Y3 = G H, T3 = E H, Z3 = F G.
template <class F, class M = MulInline>
KZ_DEV EdExt<F> ed_dbl(const EdExt<F>& p) {
    const Fe<F> A = M::template sqr<F>(p.X);
    const Fe<F> B = M::template sqr<F>(p.Y);
    const Fe<F> C = fe_dbl<F>(M::template sqr<F>(p.Z));
    const Fe<F> E = fe_sub<F>(fe_sub<F>(M::template sqr<F>(fe_add<F>(p.X, p.Y)), A), B);
    const Fe<F> D = fe_neg<F>(fe_mul5<F>(A));
    const Fe<F> G = fe_add<F>(D, B), Fv = fe_sub<F>(G, C), H = fe_sub<F>(D, B);
    EdExt<F> r;
    r.X = M::template mul<F>(E, Fv);
    r.Y = M::template mul<F>(G, H);
    r.T = M::template mul<F>(E, H);
    r.Z = M::template mul<F>(Fv, G);
    return r;
}

// Quad-cooperative forms (curve_quad.h's lane scheme: lane q computes the
// q-th product of each dependency level, a quad permute shares them).
//   addition  level 1 {X1 X2, Y1 Y2, T1 T2, Z1 Z2}, level 2 {d T1 T2,
//             (X1 + Y1)(X2 + Y2)}, level 3 {X3, Y3, T3, Z3}
//   doubling  level 1 {X^2, Y^2, Z^2, (X + Y)^2}, level 2 {X3, Y3, T3, Z3}
template <class F, class M = MulCall, class Lanes = DeviceQuad>
__host__ __device__ EdExt<F> ed_add_quad(const EdExt<F>& p, const EdExt<F>& q) {
    const int l = Lanes::lane();
    Fe<F> m = M::template mul<F>(quad_pick<F>(l, p.X, p.Y, p.T, p.Z), quad_pick<F>(l, q.X, q.Y, q.T, q.Z));
    const Fe<F> A = quad_read<0, F, Lanes>(m), B = quad_read<1, F, Lanes>(m);
    const Fe<F> TT = quad_read<2, F, Lanes>(m), D = quad_read<3, F, Lanes>(m);
    const Fe<F> S1 = fe_add<F>(p.X, p.Y), S2 = fe_add<F>(q.X, q.Y);
    m = M::template mul<F>(quad_pick<F>(l, TT, S1, TT, S1), quad_pick<F>(l, fe_const<F>(EdConst::D), S2,
                                                                            fe_const<F>(EdConst::D), S2));
    const Fe<F> C = quad_read<0, F, Lanes>(m);
    const Fe<F> E = fe_sub<F>(fe_sub<F>(quad_read<1, F, Lanes>(m), A), B);
    const Fe<F> Fv = fe_sub<F>(D, C), G = fe_add<F>(D, C), H = fe_add<F>(B, fe_mul5<F>(A));
    m = M::template mul<F>(quad_pick<F>(l, E, G, E, Fv), quad_pick<F>(l, Fv, H, H, G));
    EdExt<F> r;
    r.X = quad_read<0, F, Lanes>(m);
    r.Y = quad_read<1, F, Lanes>(m);
    r.T = quad_read<2, F, Lanes>(m);
    r.Z = quad_read<3, F, Lanes>(m);
    return r;
}

template <class F, class M = MulCall, class Lanes = DeviceQuad>
__host__ __device__ EdExt<F> ed_dbl_quad(const EdExt<F>& p) {
    const int l = Lanes::lane();
    const Fe<F> S = fe_add<F>(p.X, p.Y);
    Fe<F> m = M::template mul<F>(quad_pick<F>(l, p.X, p.Y, p.Z, S), quad_pick<F>(l, p.X, p.Y, p.Z, S));
    const Fe<F> A = quad_read<0, F, Lanes>(m), B = quad_read<1, F, Lanes>(m);
    const Fe<F> C = fe_dbl<F>(quad_read<2, F, Lanes>(m));
    const Fe<F> E = fe_sub<F>(fe_sub<F>(quad_read<3, F, Lanes>(m), A), B);
    const Fe<F> D = fe_neg<F>(fe_mul5<F>(A));
    const Fe<F> G = fe_add<F>(D, B), Fv = fe_sub<F>(G, C), H = fe_sub<F>(D, B);
    m = M::template mul<F>(quad_pick<F>(l, E, G, E, Fv), quad_pick<F>(l, Fv, H, H, G));
    EdExt<F> r;
    r.X = quad_read<0, F, Lanes>(m);
    r.Y = quad_read<1, F, Lanes>(m);
    r.T = quad_read<2, F, Lanes>(m);
    r.Z = quad_read<3, F, Lanes>(m);
    return r;
}

// Lazily reduced quad forms for chains of point operations (bucket reduction,
// window combine): inputs and outputs have coordinates below 2q.  MC is a
// canonical product policy, ML its form without the final subtraction; the
// first level's products are canonical (below q), the later ones lazy, and
// differences are one-pass a + k q - b with k q above b's bound:
//   addition  A, B, T1 T2, D < q; C < 2q; E' = (X1+Y1)(X2+Y2) < 2q
//             (operands below 4q); E = E' + 2q - A - B < 4q,
//             F = D + 2q - C < 3q, G = D + C < 3q, H = B + 5A < 6q;
//             level 3 operand pairs at most 24 q^2
//   doubling  A = X^2, B = Y^2, Z^2, S^2 < q (S = X + Y < 4q); C = 2 Z^2 < 2q;
//             D = 5q - 5A in (0, 5q]; E = S^2 + 2q - A - B < 3q; G = D + B < 6q;
//             F = G + 2q - C < 8q; H = D + q - B < 6q; level 2 pairs at most 48 q^2
// so every lazy product has operands below R q / 70.7 and stays below 2q.
template <class F, class MC, class ML, class Lanes = DeviceQuad>
__host__ __device__ EdExt<F> ed_add_quad_lazy(const EdExt<F>& p, const EdExt<F>& q) {
    const int l = Lanes::lane();
    Fe<F> m = MC::template mul<F>(quad_pick<F>(l, p.X, p.Y, p.T, p.Z), quad_pick<F>(l, q.X, q.Y, q.T, q.Z));
    const Fe<F> A = quad_read<0, F, Lanes>(m), B = quad_read<1, F, Lanes>(m);
    const Fe<F> TT = quad_read<2, F, Lanes>(m), D = quad_read<3, F, Lanes>(m);
    const Fe<F> S1 = fe_add_nored<F>(p.X, p.Y), S2 = fe_add_nored<F>(q.X, q.Y);
    m = ML::template mul<F>(quad_pick<F>(l, TT, S1, TT, S1), quad_pick<F>(l, fe_const<F>(EdConst::D), S2,
                                                                             fe_const<F>(EdConst::D), S2));
    const Fe<F> C = quad_read<0, F, Lanes>(m);
    const Fe<F> E = fe_kp_sub2<F, 2>(quad_read<1, F, Lanes>(m), A, B);
    const Fe<F> Fv = fe_sub_kp<F, 2>(D, C), G = fe_add_nored<F>(D, C);
    const Fe<F> H = fe_add_nored<F>(B, fe_small_mul<F, 5>(A));
    m = ML::template mul<F>(quad_pick<F>(l, E, G, E, Fv), quad_pick<F>(l, Fv, H, H, G));
    EdExt<F> r;
    r.X = quad_read<0, F, Lanes>(m);
    r.Y = quad_read<1, F, Lanes>(m);
    r.T = quad_read<2, F, Lanes>(m);
    r.Z = quad_read<3, F, Lanes>(m);
    return r;
}

template <class F, class MC, class ML, class Lanes = DeviceQuad>
__host__ __device__ EdExt<F> ed_dbl_quad_lazy(const EdExt<F>& p) {
    const int l = Lanes::lane();
    const Fe<F> S = fe_add_nored<F>(p.X, p.Y);
    Fe<F> m = MC::template mul<F>(quad_pick<F>(l, p.X, p.Y, p.Z, S), quad_pick<F>(l, p.X, p.Y, p.Z, S));
    const Fe<F> A = quad_read<0, F, Lanes>(m), B = quad_read<1, F, Lanes>(m);
    const Fe<F> C = fe_small_mul<F, 2>(quad_read<2, F, Lanes>(m));
    const Fe<F> E = fe_kp_sub2<F, 2>(quad_read<3, F, Lanes>(m), A, B);
    const Fe<F> D = fe_sub_kp<F, 5>(fe_zero<F>(), fe_small_mul<F, 5>(A));
    const Fe<F> G = fe_add_nored<F>(D, B), Fv = fe_sub_kp<F, 2>(G, C), H = fe_sub_kp<F, 1>(D, B);
    m = ML::template mul<F>(quad_pick<F>(l, E, G, E, Fv), quad_pick<F>(l, Fv, H, H, G));
    EdExt<F> r;
    r.X = quad_read<0, F, Lanes>(m);
    r.Y = quad_read<1, F, Lanes>(m);
    r.T = quad_read<2, F, Lanes>(m);
    r.Z = quad_read<3, F, Lanes>(m);
    return r;
}

// Affine (x, y, d x y) with one inversion; inf = the identity.
template <class F>
KZ_DEV void ed_to_affine(const EdExt<F>& p, EdAff<F>& a, bool& inf) {
    inf = ed_is_identity<F>(p);
    const Fe<F> zi = fe_inv<F>(p.Z);  // Z != 0: the formulas are complete
    a.x = fe_mul_call<F>(p.X, zi);
    a.y = fe_mul_call<F>(p.Y, zi);
    a.t = fe_mul_call<F>(fe_mul_call<F>(a.x, a.y), fe_const<F>(EdConst::D));
}

// ------------------------------------------------------------ memory I/O
// Affine records x[9] y[9] t[9] (27 words + 1 pad) in 32-word (128-byte)
// slots, one cache line per gather; extended points X Y T Z, 36 words.
struct EdRec {
    static constexpr int kAffWords = 28;
    static constexpr int kAffStride = 32;
    static constexpr int kPtWords = 36;
};

template <class F>
KZ_DEV EdAff<F> ed_aff_load(const uint32_t* __restrict__ base, size_t idx) {
    uint32_t w[EdRec::kAffWords];
    load_words<EdRec::kAffWords>(base + idx * EdRec::kAffStride, w);
    EdAff<F> a;
#pragma unroll
    for (int i = 0; i < F::N; ++i) {
        a.x.v[i] = w[i];
        a.y.v[i] = w[F::N + i];
        a.t.v[i] = w[2 * F::N + i];
    }
    return a;
}

template <class F>
KZ_DEV void ed_aff_store(uint32_t* __restrict__ base, size_t idx, const EdAff<F>& a) {
    uint32_t w[EdRec::kAffStride];
#pragma unroll
    for (int i = 0; i < EdRec::kAffStride; ++i) w[i] = 0;
#pragma unroll
    for (int i = 0; i < F::N; ++i) {
        w[i] = a.x.v[i];
        w[F::N + i] = a.y.v[i];
        w[2 * F::N + i] = a.t.v[i];
    }
    store_words<EdRec::kAffStride>(base + idx * EdRec::kAffStride, w);
}

template <class F>
KZ_DEV EdExt<F> ed_load(const uint32_t* __restrict__ base, size_t idx) {
    uint32_t w[EdRec::kPtWords];
    load_words<EdRec::kPtWords>(base + idx * EdRec::kPtWords, w);
    EdExt<F> r;
#pragma unroll
    for (int i = 0; i < F::N; ++i) {
        r.X.v[i] = w[i];
        r.Y.v[i] = w[F::N + i];
        r.T.v[i] = w[2 * F::N + i];
        r.Z.v[i] = w[3 * F::N + i];
    }
    return r;
}

template <class F>
KZ_DEV void ed_store(uint32_t* __restrict__ base, size_t idx, const EdExt<F>& p) {
    uint32_t w[EdRec::kPtWords];
#pragma unroll
    for (int i = 0; i < F::N; ++i) {
        w[i] = p.X.v[i];
        w[F::N + i] = p.Y.v[i];
        w[2 * F::N + i] = p.T.v[i];
        w[3 * F::N + i] = p.Z.v[i];
    }
    store_words<EdRec::kPtWords>(base + idx * EdRec::kPtWords, w);
}

}  // namespace kz
