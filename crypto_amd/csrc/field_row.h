// Row-distributed field arithmetic and XYZZ point operations for the serial
// chain of the window combine.
//
// The combine is one chain of ~c (W - 1) doublings per MSM: a single wave's
// instruction stream, so its time is (products on the critical path) x (one
// product's issue time).  A product in one lane (field.h) issues 2 N^2
// multiply-adds back to back.  Here the product is spread over the 16 lanes
// of a row: lane j of a row holds limb j of a field element (lanes N..15
// hold 0), and the four rows of the wave evaluate the up-to-four independent
// products of one level of the point formula (the quad-cooperative scheme of
// curve_quad.h, with rows in place of lanes).  One product then issues about
// 4 N multiply-adds per lane:
//
//   a b:  step i = 0..N-1, a_i broadcast along the row (DPP row_newbcast),
//         b rotated one lane (row_ror); lane j adds a_i b_(j-i) to column j
//         (j >= i) or a_i b_(j-i+16) to column j + 16 (j < i).  Each lane
//         keeps two 64-bit column sums; columns run 0 .. 2N - 2 <= 26.
//   reduction: step k = 0..N-1, column k (lane k) and the carry give the
//         Montgomery digit m_k = (col_k + carry) (-p^-1) mod 2^29, known to
//         the whole row; every lane adds m_k p_(j-k) to its columns, and
//         (col_k + carry + m_k p_0) >> 29 is the next carry.  The digits are
//         those of the one-lane product scan, so the result is the same
//         value, limb for limb: fe_mul_k<F, K, false>'s (below 2p, no final
//         subtraction).
//   output: columns N .. 2N - 2 rotated to lanes 0 .. N - 2, the last carry
//         added to lane 0, then carries moved up one lane per round (row_shr)
//         until every limb is below 2^29 (a wave-uniform loop on a ballot;
//         the value is below 2^(29 N), so the top limb never carries out).
//
// Sums and differences (a + k p - b with signed per-lane carries) are
// normalised by the same rounds, so every element a chain holds has limbs
// below 2^29 and the integer value the one-lane lazy code holds.  The point
// operations follow curve_quad.h's lazy formulas term for term (same bounds:
// X < 8p, Y < 4p, ZZ, ZZZ < 2p), so a chain ends in the same point with the
// same coordinates.
//
// The lane operations come from a policy W: WaveDev below on the device;
// the host check emulates a wave with the same interface.
#pragma once

#include <type_traits>

#include "curve.h"
#include "curve_ed.h"

namespace kz {

template <int I, int N, class Fn>
__host__ __device__ __forceinline__ void static_for(Fn&& fn) {
    if constexpr (I < N) {
        fn(std::integral_constant<int, I>{});
        static_for<I + 1, N>(fn);
    }
}

// One wave of 64 lanes in a workgroup of 64 threads.  DPP row controls:
// 0x110 + s row_shr:s (lane j reads lane j - s of its row; bound_ctrl makes
// it 0 below the row), 0x120 + s row_ror:s (lane j reads lane (j - s) mod
// 16), 0x150 + k row_newbcast:k (every lane reads lane k of its row).  Every
// lane gets a source or the bound_ctrl zero, so the moves need no old value.
struct WaveDev {
    using V = uint32_t;
    using V64 = uint64_t;
    using B = bool;
    __device__ static V lane16() { return threadIdx.x & 15u; }
    __device__ static V row() { return (threadIdx.x >> 4) & 3u; }
    __device__ static V c(uint32_t x) { return x; }
    __device__ static V add(V a, V b) { return a + b; }
    __device__ static V sub(V a, V b) { return a - b; }
    __device__ static V and_(V a, uint32_t m) { return a & m; }
    __device__ static V shl(V a, int s) { return a << s; }
    __device__ static V mulc(V a, uint32_t k) { return a * k; }
    __device__ static V sra(V a, int s) { return (uint32_t)((int32_t)a >> s); }
    __device__ static V shr32(V a, int s) { return a >> s; }
    __device__ static V64 wide(V a) { return a; }
    __device__ static V lo(V64 a) { return (uint32_t)a; }
    __device__ static V hi(V64 a) { return (uint32_t)(a >> 32); }
    __device__ static V64 mk64(V l, V h) { return ((uint64_t)h << 32) | l; }
    __device__ static V64 add64(V64 a, V64 b) { return a + b; }
    __device__ static V64 shr64(V64 a, int s) { return a >> s; }
    __device__ static V64 mad(V a, V b, V64 c) { return (uint64_t)a * b + c; }
    __device__ static V64 madc(V a, uint32_t b, V64 c) { return (uint64_t)a * b + c; }
    __device__ static B eqc(V a, uint32_t k) { return a == k; }
    __device__ static B ltc(V a, uint32_t k) { return a < k; }
    __device__ static B gtc(V a, uint32_t k) { return a > k; }
    __device__ static B ne(V a, V b) { return a != b; }
    __device__ static B or_(B a, B b) { return a || b; }
    __device__ static B ge(V a, V b) { return a >= b; }
    __device__ static B gt(V a, V b) { return a > b; }
    __device__ static V sel(B b, V x, V y) { return b ? x : y; }
    __device__ static V64 sel64(B b, V64 x, V64 y) { return b ? x : y; }
    template <int K>
    __device__ static V bcast(V x) {
        return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x150 + K, 0xF, 0xF, true);
    }
    template <int S>
    __device__ static V ror(V x) {
        return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x120 + S, 0xF, 0xF, true);
    }
    template <int S>
    __device__ static V shr(V x) {  // zero fill
        return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x110 + S, 0xF, 0xF, true);
    }
    // the same lane of row K
    template <int K>
    __device__ static V rowread(V x) {
        return (uint32_t)__builtin_amdgcn_ds_bpermute((int)(((K << 4) | (threadIdx.x & 15u)) << 2), (int)x);
    }
    __device__ static bool any(B b) { return __builtin_amdgcn_ballot_w64(b) != 0; }
    __device__ static uint64_t ballot(B b) { return __builtin_amdgcn_ballot_w64(b); }
    // the 16 bits of a wave mask that belong to this lane's row
    __device__ static V rowbits(uint64_t m) { return (uint32_t)(m >> (threadIdx.x & 48u)) & 0xFFFFu; }
    __device__ static uint32_t readlane(V x, int l) { return (uint32_t)__builtin_amdgcn_readlane((int)x, l); }
};

template <class W, class F>
struct FeRow {
    typename W::V v;
};

template <class W, class F>
struct XyzzRow {
    FeRow<W, F> X, Y, ZZ, ZZZ;
};

// Per-wave constants: p, k p for the differences, and the reduction's limbs
// of p per step k (lane j: p_(j-k) for j > k, p_(j-k+16) for j < k, else 0).
template <class W, class F>
struct RowConsts {
    typename W::V p, kp2, kp4, kp5, kp6, kp8;
    typename W::V plo[F::N], phi[F::N];
};

template <class W, class F, int K>
__host__ __device__ __forceinline__ typename W::V row_limbs_kp(typename W::V j) {
    typename W::V r = W::c(0);
    static_for<0, F::N>([&](auto I) { r = W::sel(W::eqc(j, I.value), W::c(kp_limb<F, K>(I.value)), r); });
    return r;
}

template <class W, class F>
__host__ __device__ __forceinline__ RowConsts<W, F> row_consts() {
    using V = typename W::V;
    RowConsts<W, F> K;
    const V j = W::lane16();
    K.p = row_limbs_kp<W, F, 1>(j);
    K.kp2 = row_limbs_kp<W, F, 2>(j);
    K.kp4 = row_limbs_kp<W, F, 4>(j);
    K.kp5 = row_limbs_kp<W, F, 5>(j);
    K.kp6 = row_limbs_kp<W, F, 6>(j);
    K.kp8 = row_limbs_kp<W, F, 8>(j);
    V rp = K.p;
    static_for<0, F::N>([&](auto I) {
        constexpr int k = I.value;
        if constexpr (k > 0) rp = W::template ror<1>(rp);
        K.plo[k] = W::sel(W::gtc(j, k), rp, W::c(0));
        K.phi[k] = W::sel(W::ltc(j, k), rp, W::c(0));
    });
    return K;
}

// Limbs held as 64-bit column sums (lane j: weight 2^(29 j), each below 2^63)
// of a value below 2^(29 N) -> normalised limbs.  After one round a limb is
// below 2^29 + 2^34 and after two below 2^29 + 2^6, so from the third round on
// the carries are single bits; those rounds repeat until none is left.
template <class W>
__host__ __device__ __forceinline__ typename W::V row_carry64(typename W::V64 r) {
    using V = typename W::V;
    const typename W::V64 c = W::shr64(r, kLimbBits);
    r = W::add64(W::wide(W::and_(W::lo(r), kLimbMask)),
                 W::mk64(W::template shr<1>(W::lo(c)), W::template shr<1>(W::hi(c))));
    V t = W::add(W::and_(W::lo(r), kLimbMask), W::template shr<1>(W::lo(W::shr64(r, kLimbBits))));
    while (W::any(W::gtc(t, kLimbMask))) t = W::add(W::and_(t, kLimbMask), W::template shr<1>(W::sra(t, kLimbBits)));
    return t;
}

// Signed 32-bit limbs (|t| < 2^31) of a value in [0, 2^(29 N)) -> normalised
// limbs.  A borrow may run above lane N - 1 before the carries from below
// arrive; the rounds then continue until it is paid back (the row's lanes are
// a value mod 2^(29 16), and the value itself is below 2^(29 N)).
template <class W>
__host__ __device__ __forceinline__ typename W::V row_carry32s(typename W::V t) {
    t = W::add(W::and_(t, kLimbMask), W::template shr<1>(W::sra(t, kLimbBits)));
    while (W::any(W::gtc(t, kLimbMask)))  // negative limbs read as large
        t = W::add(W::and_(t, kLimbMask), W::template shr<1>(W::sra(t, kLimbBits)));
    return t;
}

// Non-negative 32-bit limbs (below 2^32) -> normalised limbs.
template <class W>
__host__ __device__ __forceinline__ typename W::V row_carry32u(typename W::V t) {
    t = W::add(W::and_(t, kLimbMask), W::template shr<1>(W::shr32(t, kLimbBits)));
    while (W::any(W::gtc(t, kLimbMask))) t = W::add(W::and_(t, kLimbMask), W::template shr<1>(W::shr32(t, kLimbBits)));
    return t;
}

// a b 2^(-29 N) mod p, below p + a b / 2^(29 N): fe_mul_k<F, K, false>'s value
// and limbs, for normalised a, b with a b < 2^(29 N) p.
template <class W, class F>
__host__ __device__ __forceinline__ FeRow<W, F> row_mul(const FeRow<W, F>& a, const FeRow<W, F>& b,
                                                        const RowConsts<W, F>& K) {
    constexpr int N = F::N;
    static_assert(N <= 14, "two 64-bit column sums per lane cover 2N - 1 <= 27 columns");
    using V = typename W::V;
    using V64 = typename W::V64;
    const V j = W::lane16();
    V64 lo = W::wide(W::c(0)), hi = lo;
    V rb = b.v;
    static_for<0, N>([&](auto I) {
        constexpr int i = I.value;
        if constexpr (i > 0) rb = W::template ror<1>(rb);  // lane j: b_((j - i) mod 16)
        const V x = W::template bcast<i>(a.v);
        lo = W::mad(x, W::sel(W::ltc(j, i), W::c(0), rb), lo);
        // lanes j < i hold b_(j - i + 16), zero unless j - i + 16 < N
        if constexpr (i > 16 - N) hi = W::mad(x, W::sel(W::ltc(j, i), rb, W::c(0)), hi);
    });
    V64 carry = W::wide(W::c(0));
    static_for<0, N>([&](auto I) {
        constexpr int k = I.value;
        const V64 col = W::add64(W::mk64(W::template bcast<k>(W::lo(lo)), W::template bcast<k>(W::hi(lo))), carry);
        const V m = W::and_(W::mulc(W::lo(col), F::PINV), kLimbMask);
        carry = W::shr64(W::madc(m, F::P[0], col), kLimbBits);
        lo = W::mad(m, K.plo[k], lo);
        if constexpr (k > 16 - N) hi = W::mad(m, K.phi[k], hi);  // else phi[k] is 0 in every lane
    });
    // column c < 16 sits in lane c (lo), column c >= 16 in lane c - 16 (hi);
    // lane i takes column N + i (a rotation by 16 - N), lane 0 the last carry
    V64 r = W::sel64(W::ltc(j, N), hi, lo);
    r = W::mk64(W::template ror<16 - N>(W::lo(r)), W::template ror<16 - N>(W::hi(r)));
    r = W::add64(r, W::sel64(W::eqc(j, 0), carry, W::wide(W::c(0))));
    return FeRow<W, F>{row_carry64<W>(r)};
}

// Lazy sums and differences, the values of field.h's helpers.
template <class W, class F>
__host__ __device__ __forceinline__ FeRow<W, F> row_small_mul2(const FeRow<W, F>& a) {
    return FeRow<W, F>{row_carry32s<W>(W::shl(a.v, 1))};
}
template <class W, class F>
__host__ __device__ __forceinline__ FeRow<W, F> row_small_mul3(const FeRow<W, F>& a) {
    return FeRow<W, F>{row_carry32s<W>(W::add(W::shl(a.v, 1), a.v))};
}
// a + k p - b
template <class W, class F>
__host__ __device__ __forceinline__ FeRow<W, F> row_sub_kp(const FeRow<W, F>& a, typename W::V kp,
                                                           const FeRow<W, F>& b) {
    return FeRow<W, F>{row_carry32s<W>(W::sub(W::add(a.v, kp), b.v))};
}
// a + k p - 2b
template <class W, class F>
__host__ __device__ __forceinline__ FeRow<W, F> row_kp_2b(const FeRow<W, F>& a, typename W::V kp,
                                                          const FeRow<W, F>& b) {
    return FeRow<W, F>{row_carry32s<W>(W::sub(W::add(a.v, kp), W::shl(b.v, 1)))};
}
// r + 6p - s - 2q
template <class W, class F>
__host__ __device__ __forceinline__ FeRow<W, F> row_r6p_s_2q(const FeRow<W, F>& r, const FeRow<W, F>& s,
                                                             const FeRow<W, F>& q, const RowConsts<W, F>& K) {
    return FeRow<W, F>{row_carry32s<W>(W::sub(W::sub(W::add(r.v, K.kp6), s.v), W::shl(q.v, 1)))};
}

// Wave-uniform tests; the rows of a chain hold the same element.
template <class W, class F>
__host__ __device__ __forceinline__ bool row_is_zero(const FeRow<W, F>& a) {
    return !W::any(W::ne(a.v, W::c(0)));
}
template <class W, class F>
__host__ __device__ __forceinline__ bool row_is_zero_lazy(const FeRow<W, F>& a, const RowConsts<W, F>& K) {
    return row_is_zero<W, F>(a) || !W::any(W::ne(a.v, K.p));
}

template <class W, class F>
__host__ __device__ __forceinline__ FeRow<W, F> row_pick(typename W::V q, const FeRow<W, F>& a0,
                                                         const FeRow<W, F>& a1, const FeRow<W, F>& a2,
                                                         const FeRow<W, F>& a3) {
    return FeRow<W, F>{W::sel(W::eqc(q, 0), a0.v, W::sel(W::eqc(q, 1), a1.v, W::sel(W::eqc(q, 2), a2.v, a3.v)))};
}

template <int K, class W, class F>
__host__ __device__ __forceinline__ FeRow<W, F> row_read(const FeRow<W, F>& a) {
    return FeRow<W, F>{W::template rowread<K>(a.v)};
}

template <class W, class F>
__host__ __device__ __forceinline__ bool row_pt_is_inf(const XyzzRow<W, F>& p) {
    return row_is_zero<W, F>(p.ZZ);
}

// 2P, three product levels (xyzz_dbl_quad under a lazy policy).
template <class W, class F>
__host__ __device__ XyzzRow<W, F> row_dbl(const XyzzRow<W, F>& p, const RowConsts<W, F>& K) {
    using E = FeRow<W, F>;
    const typename W::V q = W::row();
    const E U = row_small_mul2<W, F>(p.Y);
    E m = row_mul<W, F>(row_pick<W, F>(q, U, p.X, U, p.X), row_pick<W, F>(q, U, p.X, U, p.X), K);
    const E V = row_read<0, W, F>(m), XX = row_read<1, W, F>(m);
    const E X3 = row_small_mul3<W, F>(XX);
    m = row_mul<W, F>(row_pick<W, F>(q, U, p.X, V, X3), row_pick<W, F>(q, V, V, p.ZZ, X3), K);
    const E Wt = row_read<0, W, F>(m), S = row_read<1, W, F>(m);
    XyzzRow<W, F> r;
    r.ZZ = row_read<2, W, F>(m);
    r.X = row_kp_2b<W, F>(row_read<3, W, F>(m), K.kp4, S);
    const E SmX = row_sub_kp<W, F>(S, K.kp6, r.X);
    m = row_mul<W, F>(row_pick<W, F>(q, X3, Wt, Wt, Wt), row_pick<W, F>(q, SmX, p.Y, p.ZZZ, p.ZZZ), K);
    r.Y = row_sub_kp<W, F>(row_read<0, W, F>(m), K.kp2, row_read<1, W, F>(m));
    r.ZZZ = row_read<2, W, F>(m);
    return r;
}

// a + b with the doubling and inverse cases, four product levels
// (xyzz_add_quad under a lazy policy).
template <class W, class F>
__host__ __device__ XyzzRow<W, F> row_add(const XyzzRow<W, F>& a, const XyzzRow<W, F>& b,
                                          const RowConsts<W, F>& K) {
    using E = FeRow<W, F>;
    if (row_pt_is_inf<W, F>(a)) return b;
    if (row_pt_is_inf<W, F>(b)) return a;
    const typename W::V q = W::row();
    E m = row_mul<W, F>(row_pick<W, F>(q, a.X, b.X, a.Y, b.Y), row_pick<W, F>(q, b.ZZ, a.ZZ, b.ZZZ, a.ZZZ), K);
    const E U1 = row_read<0, W, F>(m), S1 = row_read<2, W, F>(m);
    const E P = row_sub_kp<W, F>(row_read<1, W, F>(m), K.kp2, U1);
    const E R = row_sub_kp<W, F>(row_read<3, W, F>(m), K.kp2, S1);
    m = row_mul<W, F>(row_pick<W, F>(q, P, R, a.ZZ, a.ZZZ), row_pick<W, F>(q, P, R, b.ZZ, b.ZZZ), K);
    const E PP = row_read<0, W, F>(m), RR = row_read<1, W, F>(m), Z12 = row_read<2, W, F>(m),
            ZZZ12 = row_read<3, W, F>(m);
    if (row_is_zero_lazy<W, F>(PP, K)) {
        if (row_is_zero_lazy<W, F>(RR, K)) return row_dbl<W, F>(a, K);
        XyzzRow<W, F> z;
        z.X = z.Y = z.ZZ = z.ZZZ = E{W::c(0)};
        return z;
    }
    m = row_mul<W, F>(row_pick<W, F>(q, P, U1, Z12, Z12), row_pick<W, F>(q, PP, PP, PP, PP), K);
    const E PPP = row_read<0, W, F>(m), Q = row_read<1, W, F>(m);
    XyzzRow<W, F> r;
    r.ZZ = row_read<2, W, F>(m);
    r.X = row_r6p_s_2q<W, F>(RR, PPP, Q, K);
    const E QmX = row_sub_kp<W, F>(Q, K.kp8, r.X);
    m = row_mul<W, F>(row_pick<W, F>(q, R, S1, ZZZ12, ZZZ12), row_pick<W, F>(q, QmX, PPP, PPP, PPP), K);
    r.Y = row_sub_kp<W, F>(row_read<0, W, F>(m), K.kp2, row_read<1, W, F>(m));
    r.ZZZ = row_read<2, W, F>(m);
    return r;
}

// a mod p for a normalised a < 2p: a - p when a >= p.  Per row, a >= p when
// the highest limb where a and p differ is larger in a (or none differs):
// with G, L the row's lanes where a's limb is above / below p's, a >= p
// exactly when G >= L as 16-bit integers (the sets are disjoint).
template <class W, class F>
__host__ __device__ __forceinline__ FeRow<W, F> row_reduce_once(const FeRow<W, F>& a, const RowConsts<W, F>& K) {
    const typename W::V G = W::rowbits(W::ballot(W::gt(a.v, K.p)));
    const typename W::V L = W::rowbits(W::ballot(W::gt(K.p, a.v)));
    return FeRow<W, F>{row_carry32s<W>(W::sel(W::ge(G, L), W::sub(a.v, K.p), a.v))};
}

// a b 2^(-29 N) mod p, canonical: fe_mul_k<F, K, true>'s value and limbs.
template <class W, class F>
__host__ __device__ __forceinline__ FeRow<W, F> row_mul_canon(const FeRow<W, F>& a, const FeRow<W, F>& b,
                                                              const RowConsts<W, F>& K) {
    return row_reduce_once<W, F>(row_mul<W, F>(a, b, K), K);
}

// a + b, no reduction (fe_add_nored)
template <class W, class F>
__host__ __device__ __forceinline__ FeRow<W, F> row_add_nored(const FeRow<W, F>& a, const FeRow<W, F>& b) {
    return FeRow<W, F>{row_carry32u<W>(W::add(a.v, b.v))};
}
// a + k p - b - c (fe_kp_sub2)
template <class W, class F>
__host__ __device__ __forceinline__ FeRow<W, F> row_kp_sub2(const FeRow<W, F>& a, typename W::V kp,
                                                            const FeRow<W, F>& b, const FeRow<W, F>& c) {
    return FeRow<W, F>{row_carry32s<W>(W::sub(W::sub(W::add(a.v, kp), b.v), c.v))};
}
template <class W, class F>
__host__ __device__ __forceinline__ FeRow<W, F> row_small_mul5(const FeRow<W, F>& a) {
    return FeRow<W, F>{row_carry32u<W>(W::add(W::shl(a.v, 2), a.v))};
}

// Extended twisted Edwards points (Banderwagon), by rows.
template <class W, class F>
struct EdRow {
    FeRow<W, F> X, Y, T, Z;
};

template <class W, class F>
__host__ __device__ __forceinline__ bool row_ed_is_identity(const EdRow<W, F>& p) {
    return row_is_zero<W, F>(p.X) && !W::any(W::ne(p.Y.v, p.Z.v));
}

// p + q (ed_add_quad_lazy with a canonical first level): three product levels.
template <class W, class F>
__host__ __device__ EdRow<W, F> row_ed_add(const EdRow<W, F>& p, const EdRow<W, F>& q, const FeRow<W, F>& d,
                                           const RowConsts<W, F>& K) {
    using E = FeRow<W, F>;
    const typename W::V l = W::row();
    E m = row_mul_canon<W, F>(row_pick<W, F>(l, p.X, p.Y, p.T, p.Z), row_pick<W, F>(l, q.X, q.Y, q.T, q.Z), K);
    const E A = row_read<0, W, F>(m), B = row_read<1, W, F>(m), TT = row_read<2, W, F>(m), D = row_read<3, W, F>(m);
    const E S1 = row_add_nored<W, F>(p.X, p.Y), S2 = row_add_nored<W, F>(q.X, q.Y);
    m = row_mul<W, F>(row_pick<W, F>(l, TT, S1, TT, S1), row_pick<W, F>(l, d, S2, d, S2), K);
    const E C = row_read<0, W, F>(m);
    const E Ev = row_kp_sub2<W, F>(row_read<1, W, F>(m), K.kp2, A, B);
    const E Fv = row_sub_kp<W, F>(D, K.kp2, C), G = row_add_nored<W, F>(D, C);
    const E H = row_add_nored<W, F>(B, row_small_mul5<W, F>(A));
    m = row_mul<W, F>(row_pick<W, F>(l, Ev, G, Ev, Fv), row_pick<W, F>(l, Fv, H, H, G), K);
    return EdRow<W, F>{row_read<0, W, F>(m), row_read<1, W, F>(m), row_read<2, W, F>(m), row_read<3, W, F>(m)};
}

// 2p (ed_dbl_quad_lazy with a canonical first level): two product levels.
template <class W, class F>
__host__ __device__ EdRow<W, F> row_ed_dbl(const EdRow<W, F>& p, const RowConsts<W, F>& K) {
    using E = FeRow<W, F>;
    const typename W::V l = W::row();
    const E S = row_add_nored<W, F>(p.X, p.Y);
    E m = row_mul_canon<W, F>(row_pick<W, F>(l, p.X, p.Y, p.Z, S), row_pick<W, F>(l, p.X, p.Y, p.Z, S), K);
    const E A = row_read<0, W, F>(m), B = row_read<1, W, F>(m);
    const E C = row_small_mul2<W, F>(row_read<2, W, F>(m));
    const E Ev = row_kp_sub2<W, F>(row_read<3, W, F>(m), K.kp2, A, B);
    const E D = row_sub_kp<W, F>(E{W::c(0)}, K.kp5, row_small_mul5<W, F>(A));
    const E G = row_add_nored<W, F>(D, B), Fv = row_sub_kp<W, F>(G, K.kp2, C), H = row_sub_kp<W, F>(D, K.p, B);
    m = row_mul<W, F>(row_pick<W, F>(l, Ev, G, Ev, Fv), row_pick<W, F>(l, Fv, H, H, G), K);
    return EdRow<W, F>{row_read<0, W, F>(m), row_read<1, W, F>(m), row_read<2, W, F>(m), row_read<3, W, F>(m)};
}

template <class W, class F>
__host__ __device__ __forceinline__ EdRow<W, F> row_from_ed(const EdExt<F>& p) {
    return EdRow<W, F>{row_from_fe<W, F>(p.X), row_from_fe<W, F>(p.Y), row_from_fe<W, F>(p.T),
                       row_from_fe<W, F>(p.Z)};
}

// Standard (one-lane) form <-> rows.
template <class W, class F>
__host__ __device__ __forceinline__ FeRow<W, F> row_from_fe(const Fe<F>& a) {
    const typename W::V j = W::lane16();
    typename W::V r = W::c(0);
    static_for<0, F::N>([&](auto I) { r = W::sel(W::eqc(j, I.value), W::c(a.v[I.value]), r); });
    return FeRow<W, F>{r};
}
template <class W, class F>
__host__ __device__ __forceinline__ Fe<F> row_to_fe(const FeRow<W, F>& a) {
    Fe<F> r;
    static_for<0, F::N>([&](auto I) { r.v[I.value] = W::readlane(a.v, I.value); });
    return r;
}
template <class W, class F>
__host__ __device__ __forceinline__ XyzzRow<W, F> row_from_xyzz(const Xyzz<F>& p) {
    return XyzzRow<W, F>{row_from_fe<W, F>(p.X), row_from_fe<W, F>(p.Y), row_from_fe<W, F>(p.ZZ),
                         row_from_fe<W, F>(p.ZZZ)};
}
template <class W, class F>
__host__ __device__ __forceinline__ Xyzz<F> row_to_xyzz(const XyzzRow<W, F>& p) {
    Xyzz<F> r;
    r.X = row_to_fe<W, F>(p.X);
    r.Y = row_to_fe<W, F>(p.Y);
    r.ZZ = row_to_fe<W, F>(p.ZZ);
    r.ZZZ = row_to_fe<W, F>(p.ZZZ);
    return r;
}
template <class W, class F>
__host__ __device__ __forceinline__ EdExt<F> row_to_ed(const EdRow<W, F>& p) {
    EdExt<F> r;
    r.X = row_to_fe<W, F>(p.X);
    r.Y = row_to_fe<W, F>(p.Y);
    r.T = row_to_fe<W, F>(p.T);
    r.Z = row_to_fe<W, F>(p.Z);
    return r;
}

}  // namespace kz
