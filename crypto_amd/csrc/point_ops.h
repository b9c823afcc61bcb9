// The point operations the MSM kernels use, selected by the base-field type:
// XYZZ on the short-Weierstrass curves (curve.h, curve_quad.h), extended
// twisted Edwards coordinates for Banderwagon (curve_ed.h, field type BwFp).
// The kernels of msm_impl.h and gen_impl.h are written once against this
// interface:
//   Pt, Aff              accumulator / projective point, gathered affine record
//   kPtWords, kAffStride record sizes in 32-bit words
//   inf, is_inf          identity
//   from_aff, neg_aff    an affine record as a point; its negation
//   load(_aff), store(_aff)
//   madd_acc             the accumulation loop's mixed addition (its own bounds)
//   canon                canonical coordinates for a store
//   add_q<M>, dbl_q<M>   quad-cooperative addition and doubling, policy M
//   add<M>, dbl<M>, madd<M>, to_affine   serial forms (generator tables)
#pragma once

#include "curve_ed.h"

namespace kz {

#define KZ_DEVI __device__ __forceinline__

template <class F>
struct PointOps {
    using Pt = Xyzz<F>;
    using Aff = Affine<F>;
    static constexpr int kPtWords = Rec<F>::kXyzzWords;
    static constexpr int kAffStride = Rec<F>::kAffStride;
    KZ_DEVI static Pt inf() { return xyzz_inf<F>(); }
    KZ_DEVI static bool is_inf(const Pt& p) { return xyzz_is_inf<F>(p); }
    KZ_DEVI static Pt from_aff(const Aff& a) { return xyzz_from_affine<F>(a); }
    KZ_DEVI static Aff neg_aff(Aff a) {
        a.y = fe_neg<F>(a.y);
        return a;
    }
    KZ_DEVI static Aff load_aff(const uint32_t* __restrict__ b, size_t i) { return affine_load<F>(b, (uint32_t)i); }
    KZ_DEVI static void store_aff(uint32_t* __restrict__ b, size_t i, const Aff& a) { affine_store<F>(b, i, a); }
    KZ_DEVI static Pt load(const uint32_t* __restrict__ b, size_t i) { return xyzz_load<F>(b, i); }
    KZ_DEVI static void store(uint32_t* __restrict__ b, size_t i, const Pt& p) { xyzz_store<F>(b, i, p); }
    // coordinates below 8p / 4p / 2p until stored
    KZ_DEVI static void madd_acc(Pt& acc, const Aff& q) { xyzz_madd_lazy<F>(acc, q); }
    KZ_DEVI static Pt canon(const Pt& p) { return xyzz_canon_acc<F>(p); }
    template <class M>
    KZ_DEVI static Pt add_q(const Pt& a, const Pt& b) {
        return xyzz_add_quad<F, M>(a, b);
    }
    template <class M>
    KZ_DEVI static Pt dbl_q(const Pt& a) {
        return xyzz_dbl_quad<F, M>(a);
    }
    template <class M>
    KZ_DEVI static Pt add(const Pt& a, const Pt& b) {
        return xyzz_add<F, M>(a, b);
    }
    template <class M>
    KZ_DEVI static Pt dbl(const Pt& a) {
        return xyzz_dbl<F, M>(a);
    }
    template <class M>
    KZ_DEVI static void madd(Pt& acc, const Aff& q) {
        xyzz_madd<F, M>(acc, q);
    }
    KZ_DEVI static void to_affine(const Pt& p, Aff& a, bool& inf) { xyzz_to_affine<F>(p, a, inf); }
};

// The canonical product policy of a lazy one (Edwards values stay canonical).
template <class M>
struct CanonMul {
    using type = M;
};
template <>
struct CanonMul<MulLazy> {
    using type = MulInline;
};
template <>
struct CanonMul<MulCallLazy> {
    using type = MulCall;
};

template <>
struct PointOps<BwFp> {
    using F = BwFp;
    using Pt = EdExt<F>;
    using Aff = EdAff<F>;
    static constexpr int kPtWords = EdRec::kPtWords;
    static constexpr int kAffStride = EdRec::kAffStride;
    KZ_DEVI static Pt inf() { return ed_identity<F>(); }
    KZ_DEVI static bool is_inf(const Pt& p) { return ed_is_identity<F>(p); }
    KZ_DEVI static Pt from_aff(const Aff& a) { return ed_from_aff<F>(a); }
    KZ_DEVI static Aff neg_aff(const Aff& a) { return ed_neg_aff<F>(a); }
    KZ_DEVI static Aff load_aff(const uint32_t* __restrict__ b, size_t i) { return ed_aff_load<F>(b, i); }
    KZ_DEVI static void store_aff(uint32_t* __restrict__ b, size_t i, const Aff& a) { ed_aff_store<F>(b, i, a); }
    KZ_DEVI static Pt load(const uint32_t* __restrict__ b, size_t i) { return ed_load<F>(b, i); }
    KZ_DEVI static void store(uint32_t* __restrict__ b, size_t i, const Pt& p) { ed_store<F>(b, i, p); }
    // coordinates below 2q until stored
    KZ_DEVI static void madd_acc(Pt& acc, const Aff& q) { ed_madd_lazy<F>(acc, q); }
    KZ_DEVI static Pt canon(const Pt& p) { return ed_canon<F>(p); }
    // lazy policies keep coordinates below 2q (curve_ed.h); canonical ones
    // keep them canonical
    template <class M>
    KZ_DEVI static Pt add_q(const Pt& a, const Pt& b) {
        if constexpr (M::kLazy) return ed_add_quad_lazy<F, typename CanonMul<M>::type, M>(a, b);
        else return ed_add_quad<F, M>(a, b);
    }
    template <class M>
    KZ_DEVI static Pt dbl_q(const Pt& a) {
        if constexpr (M::kLazy) return ed_dbl_quad_lazy<F, typename CanonMul<M>::type, M>(a);
        else return ed_dbl_quad<F, M>(a);
    }
    template <class M>
    KZ_DEVI static Pt add(const Pt& a, const Pt& b) {
        return ed_add<F, typename CanonMul<M>::type>(a, b);
    }
    template <class M>
    KZ_DEVI static Pt dbl(const Pt& a) {
        return ed_dbl<F, typename CanonMul<M>::type>(a);
    }
    template <class M>
    KZ_DEVI static void madd(Pt& acc, const Aff& q) {
        ed_madd<F, typename CanonMul<M>::type>(acc, q);
    }
    KZ_DEVI static void to_affine(const Pt& p, Aff& a, bool& inf) { ed_to_affine<F>(p, a, inf); }
};

}  // namespace kz
