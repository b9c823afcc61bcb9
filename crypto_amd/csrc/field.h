// Finite-field arithmetic for gfx950: Montgomery form over unsaturated
// radix-2^29 limbs, host- and device-compilable (the host build is what the
// CPU tests use to check the arithmetic against the oracle).
//
// Why radix 2^29: the product-scanning Montgomery multiply below accumulates a
// whole column (up to 2N products of 29-bit limbs, < 2N * 2^58 < 2^63) in one
// 64-bit register pair, so every partial product is a single v_mad_u64_u32
// whose addend is the accumulator itself -- no carry handling inside a
// column and no register shuffling to build 64-bit addends.  With saturated
// 32-bit limbs the same multiply compiled to ~2 v_mov per v_mad.
//
// Fields (N limbs, R = 2^(29N)):
//   BLS12-381 Fp  N = 14  (381-bit p)   bls/cpp/bls12_381_fp.hpp:83-85
//   BLS12-381 Fr  N = 9   (255-bit r)   kzg/cpp/kzg_oracle.hpp:66-71
//   BN254 Fp, Fr  N = 9   (254-bit)     bn254/cpp/bn254_fp.hpp:75-81
//   secp256k1 Fp  N = 10  (256-bit p)   secp256k1/cpp/field.hpp:67
// All five satisfy 4p < R, so Montgomery products of values below p stay
// below 2p before the final conditional subtraction.
//
// Invariant: every Fe produced by the arithmetic has limbs < 2^29 and value
// < p, so representations are unique and equality is limb equality.
// Montgomery forms differ from the reference's (R = 2^384 / 2^256) but never
// leave the device: inputs and outputs are canonical integers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define KZ_DEV __host__ __device__ __forceinline__

namespace kz {

constexpr int kLimbBits = 29;
constexpr uint32_t kLimbMask = (1u << kLimbBits) - 1u;

struct BlsFp {
    static constexpr int N = 14;
    static constexpr int BITS = 381;
    static constexpr uint32_t PINV = 0x1FFCFFFDu;
    static constexpr uint32_t P[14] = {0x1FFFAAABu, 0x0FF7FFFFu, 0x14FFFFEEu, 0x17FFFD62u, 0x0F6241EAu,
                                       0x09507B58u, 0x0AFD9CC3u, 0x109E70A2u, 0x1764774Bu, 0x121A5D66u,
                                       0x12C6E9EDu, 0x12FFCD34u, 0x00111EA3u, 0x0000000Du};
    static constexpr uint32_t ONE[14] = {0x03A9FB84u, 0x0BA00690u, 0x071288F1u, 0x0F59BCC5u, 0x126CB614u,
                                         0x0585BF36u, 0x1B85AC3Du, 0x1CF856FAu, 0x1891ECBDu, 0x1A7EEC05u,
                                         0x155A88F0u, 0x0741AC6Du, 0x1317C30Fu, 0x00000009u};
    static constexpr uint32_t R2[14] = {0x15BEF7AEu, 0x1031CD0Eu, 0x02DD93E8u, 0x09226323u, 0x0E6E2CD2u,
                                        0x11684DAAu, 0x1170E5DBu, 0x088E25B1u, 0x1B366399u, 0x1C536F47u,
                                        0x0D1F9CBCu, 0x0278B67Fu, 0x1EA66A2Bu, 0x0000000Cu};
    static constexpr uint32_t R3[14] = {0x09217D6Au, 0x1D6118BAu, 0x1114B11Cu, 0x0126AEE7u, 0x0A55E2C4u,
                                        0x04D63CE0u, 0x154FF87Du, 0x14555478u, 0x1D1BDC0Du, 0x161F98D4u,
                                        0x1D74E921u, 0x09B4345Au, 0x1E5ECFB8u, 0x0000000Au};
};

struct BlsFr {
    static constexpr int N = 9;
    static constexpr int BITS = 255;
    static constexpr uint32_t PINV = 0x1FFFFFFFu;
    static constexpr uint32_t P[9] = {0x00000001u, 0x1FFFFFF8u, 0x1F96FFBFu, 0x1B4805FFu, 0x1D80553Bu,
                                      0x0C0404D0u, 0x1520CCE7u, 0x0A6533AFu, 0x0073EDA7u};
    static constexpr uint32_t ONE[9] = {0x1FFFFFBAu, 0x0000022Fu, 0x1CB61180u, 0x0A4E5C00u, 0x0EE8B1A2u,
                                        0x16E6AEDFu, 0x1907F8BBu, 0x0853DDF7u, 0x004D043Fu};
    static constexpr uint32_t R2[9] = {0x0A71B3C0u, 0x1D32207Eu, 0x1663D999u, 0x1C5ABC93u, 0x03B58C44u,
                                       0x0BE37438u, 0x0829F771u, 0x1660139Eu, 0x0027FD91u};
    static constexpr uint32_t R3[9] = {0x19D7065Du, 0x0020DB85u, 0x16122E43u, 0x0EDB1FF8u, 0x0FDA6124u,
                                       0x0517AC72u, 0x12E6A522u, 0x19D54EDAu, 0x0009750Bu};
};

struct BnFp {
    static constexpr int N = 9;
    static constexpr int BITS = 254;
    static constexpr uint32_t PINV = 0x04866389u;
    static constexpr uint32_t P[9] = {0x187CFD47u, 0x010460B6u, 0x1C72A34Fu, 0x02D522D0u, 0x1585D978u,
                                      0x02DB40C0u, 0x00A6E141u, 0x0E5C2634u, 0x0030644Eu};
    static constexpr uint32_t ONE[9] = {0x157CCC21u, 0x141C2758u, 0x185230D3u, 0x014C0419u, 0x0AA36FB9u,
                                        0x1D4240CEu, 0x11D54C07u, 0x052AC7A8u, 0x000DC836u};
    static constexpr uint32_t R2[9] = {0x059BAC10u, 0x0D1503A3u, 0x018016B8u, 0x10AB0CA8u, 0x02632639u,
                                       0x02C0169Fu, 0x169BFD53u, 0x11869D4Cu, 0x002A11A6u};
    static constexpr uint32_t R3[9] = {0x0E2312B2u, 0x16C05CA2u, 0x0BC84389u, 0x1CDF310Bu, 0x11ADAFDDu,
                                       0x032E568Eu, 0x1D6AE48Cu, 0x10D4CD1Fu, 0x0026C2D2u};
};

struct BnFr {
    static constexpr int N = 9;
    static constexpr int BITS = 254;
    static constexpr uint32_t PINV = 0x0FFFFFFFu;
    static constexpr uint32_t P[9] = {0x10000001u, 0x1F0FAC9Fu, 0x0E5C2450u, 0x07D090F3u, 0x1585D283u,
                                      0x02DB40C0u, 0x00A6E141u, 0x0E5C2634u, 0x0030644Eu};
    static constexpr uint32_t ONE[9] = {0x0FFFFF57u, 0x1EA70AB4u, 0x052C068Bu, 0x17504F49u, 0x0AA8075Bu,
                                        0x1D4240CEu, 0x11D54C07u, 0x052AC7A8u, 0x000DC836u};
    static constexpr uint32_t R2[9] = {0x05B69BD4u, 0x06170A5Au, 0x020CDDCEu, 0x1DB6310Bu, 0x0E54D0FFu,
                                       0x1CF855E3u, 0x1C15E103u, 0x07D09161u, 0x000A054Au};
    static constexpr uint32_t R3[9] = {0x001FDDB2u, 0x17D30B63u, 0x1A2600EEu, 0x09507C47u, 0x1496B29Bu,
                                       0x0B00A268u, 0x15B645EBu, 0x1F9FCB3Du, 0x001BAA96u};
};

// secp256k1 Fp, p = 2^256 - 2^32 - 977 (secp256k1/cpp/field.hpp:67).  Ten limbs
// although nine hold p: with R = 2^290, R / p = 2^34 keeps the wide operand
// bounds of the lazy point formulas (nine limbs would give R / p = 32 < 100).
struct SecpFp {
    static constexpr int N = 10;
    static constexpr int BITS = 256;
    static constexpr uint32_t PINV = 0x12253531u;
    static constexpr uint32_t P[10] = {0x1FFFFC2Fu, 0x1FFFFFF7u, 0x1FFFFFFFu, 0x1FFFFFFFu, 0x1FFFFFFFu,
                                       0x1FFFFFFFu, 0x1FFFFFFFu, 0x1FFFFFFFu, 0x00FFFFFFu, 0x00000000u};
    static constexpr uint32_t ONE[10] = {0x00000000u, 0x00007A20u, 0x00000100u, 0x00000000u, 0x00000000u,
                                         0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u};
    static constexpr uint32_t R2[10] = {0x00000000u, 0x00000000u, 0x1A428400u, 0x00F44001u, 0x00010000u,
                                        0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u};
    static constexpr uint32_t R3[10] = {0x00000000u, 0x00000000u, 0x00000000u, 0x1B388000u, 0x078CDE57u,
                                        0x0E600576u, 0x0100000Bu, 0x00000000u, 0x00000000u, 0x00000000u};
};

template <class F>
struct Fe {
    uint32_t v[F::N];
};

// ------------------------------------------------------------------ basics
template <class F>
KZ_DEV Fe<F> fe_zero() {
    Fe<F> r;
#pragma unroll
    for (int i = 0; i < F::N; ++i) r.v[i] = 0;
    return r;
}

template <class F>
KZ_DEV Fe<F> fe_const(const uint32_t (&c)[F::N]) {
    Fe<F> r;
#pragma unroll
    for (int i = 0; i < F::N; ++i) r.v[i] = c[i];
    return r;
}

template <class F>
KZ_DEV Fe<F> fe_one() {
    return fe_const<F>(F::ONE);
}

template <class F>
KZ_DEV bool fe_is_zero(const Fe<F>& a) {
    uint32_t acc = 0;
#pragma unroll
    for (int i = 0; i < F::N; ++i) acc |= a.v[i];
    return acc == 0;
}

template <class F>
KZ_DEV bool fe_eq(const Fe<F>& a, const Fe<F>& b) {
    uint32_t acc = 0;
#pragma unroll
    for (int i = 0; i < F::N; ++i) acc |= a.v[i] ^ b.v[i];
    return acc == 0;
}

// a - p if a >= p, else a; a normalised (limbs < 2^29), a < 2^(29N).
template <class F>
KZ_DEV void fe_reduce_once(Fe<F>& a) {
    uint32_t d[F::N];
    uint32_t borrow = 0;
#pragma unroll
    for (int i = 0; i < F::N; ++i) {
        const uint32_t t = a.v[i] - F::P[i] - borrow;
        borrow = t >> 31;
        d[i] = t & kLimbMask;
    }
#pragma unroll
    for (int i = 0; i < F::N; ++i) a.v[i] = borrow ? a.v[i] : d[i];
}

// Is the normalised value a < p?
template <class F>
KZ_DEV bool fe_is_canonical(const Fe<F>& a) {
    uint32_t borrow = 0;
#pragma unroll
    for (int i = 0; i < F::N; ++i) borrow = (a.v[i] - F::P[i] - borrow) >> 31;
    return borrow != 0;
}

template <class F>
KZ_DEV Fe<F> fe_add(const Fe<F>& a, const Fe<F>& b) {
    Fe<F> s;
    uint32_t carry = 0;
#pragma unroll
    for (int i = 0; i < F::N; ++i) {
        const uint32_t t = a.v[i] + b.v[i] + carry;
        carry = t >> kLimbBits;
        s.v[i] = t & kLimbMask;
    }
    // a + b < 2p < 2^(29N): no carry out of the top limb.
    fe_reduce_once<F>(s);
    return s;
}

template <class F>
KZ_DEV Fe<F> fe_sub(const Fe<F>& a, const Fe<F>& b) {
    Fe<F> d;
    uint32_t borrow = 0;
#pragma unroll
    for (int i = 0; i < F::N; ++i) {
        const uint32_t t = a.v[i] - b.v[i] - borrow;
        borrow = t >> 31;
        d.v[i] = t & kLimbMask;
    }
    // on borrow the limbs hold a - b + 2^(29N); adding p wraps back into range
    const uint32_t mask = 0u - borrow;
    uint32_t carry = 0;
#pragma unroll
    for (int i = 0; i < F::N; ++i) {
        const uint32_t t = d.v[i] + (F::P[i] & mask) + carry;
        carry = t >> kLimbBits;
        d.v[i] = t & kLimbMask;
    }
    return d;
}

template <class F>
KZ_DEV Fe<F> fe_dbl(const Fe<F>& a) {
    return fe_add<F>(a, a);
}

template <class F>
KZ_DEV Fe<F> fe_neg(const Fe<F>& a) {
    return fe_sub<F>(fe_zero<F>(), a);
}

// Montgomery product a b 2^(-29N) mod p by product scanning: column k of
// a*b + m*p accumulates in 64-bit values; for k < N the digit m_k is chosen to
// zero the column's low 29 bits.  2N^2 multiply-adds.
//
// K independent accumulators per column (summed at the column's end, K-1
// 64-bit adds) give K-way instruction-level parallelism: a single dependent
// chain of v_mad_u64_u32 stalls on the instruction's latency whenever few
// waves share a SIMD.  Each accumulator holds at most N products of 29-bit
// limbs plus the column carry, far below 2^64.
template <class F, int K, bool Canon = true>
KZ_DEV Fe<F> fe_mul_k(const Fe<F>& a, const Fe<F>& b) {
    constexpr int N = F::N;
    uint32_t m[N];
    Fe<F> t;
    uint64_t acc[K];
#pragma unroll
    for (int q = 0; q < K; ++q) acc[q] = 0;
#pragma unroll
    for (int k = 0; k < N; ++k) {
#pragma unroll
        for (int i = 0; i <= k; ++i) acc[i % K] += (uint64_t)a.v[i] * b.v[k - i];
#pragma unroll
        for (int i = 0; i < k; ++i) acc[(k + 1 + i) % K] += (uint64_t)m[i] * F::P[k - i];
        uint64_t s = acc[0];
#pragma unroll
        for (int q = 1; q < K; ++q) s += acc[q];
        m[k] = ((uint32_t)s * F::PINV) & kLimbMask;
        s += (uint64_t)m[k] * F::P[0];
        acc[0] = s >> kLimbBits;
#pragma unroll
        for (int q = 1; q < K; ++q) acc[q] = 0;
    }
#pragma unroll
    for (int k = N; k < 2 * N - 1; ++k) {
        int q = 0;
#pragma unroll
        for (int i = k - N + 1; i < N; ++i) {
            acc[q] += (uint64_t)a.v[i] * b.v[k - i];
            q = (q + 1) % K;
            acc[q] += (uint64_t)m[i] * F::P[k - i];
            q = (q + 1) % K;
        }
        uint64_t s = acc[0];
#pragma unroll
        for (int r = 1; r < K; ++r) s += acc[r];
        t.v[k - N] = (uint32_t)s & kLimbMask;
        acc[0] = s >> kLimbBits;
#pragma unroll
        for (int r = 1; r < K; ++r) acc[r] = 0;
    }
    t.v[N - 1] = (uint32_t)acc[0];  // < 2p < 2^(29N): fits the top limb
    if (Canon) fe_reduce_once<F>(t);
    return t;
}

// (a b + c d) 2^(-29N) mod p: both products scanned into one set of column
// sums, then one Montgomery reduction (N^2 fewer multiply-adds than two
// products and a sum).  No final subtraction: the result is below
// p + (a b + c d) / 2^(29N).  A column holds at most 3N - 1 terms below 2^58
// plus the carry from the column before: below 2^64 for N <= 14.
template <class F>
KZ_DEV Fe<F> fe_mul2_lazy(const Fe<F>& a, const Fe<F>& b, const Fe<F>& c, const Fe<F>& d) {
    constexpr int N = F::N;
    static_assert(3 * N - 1 <= 63, "column sums must stay below 2^64");
    uint32_t m[N];
    Fe<F> t;
    uint64_t acc = 0;
#pragma unroll
    for (int k = 0; k < N; ++k) {
#pragma unroll
        for (int i = 0; i <= k; ++i) {
            acc += (uint64_t)a.v[i] * b.v[k - i];
            acc += (uint64_t)c.v[i] * d.v[k - i];
        }
#pragma unroll
        for (int i = 0; i < k; ++i) acc += (uint64_t)m[i] * F::P[k - i];
        m[k] = ((uint32_t)acc * F::PINV) & kLimbMask;
        acc += (uint64_t)m[k] * F::P[0];
        acc >>= kLimbBits;
    }
#pragma unroll
    for (int k = N; k < 2 * N - 1; ++k) {
#pragma unroll
        for (int i = k - N + 1; i < N; ++i) {
            acc += (uint64_t)a.v[i] * b.v[k - i];
            acc += (uint64_t)c.v[i] * d.v[k - i];
            acc += (uint64_t)m[i] * F::P[k - i];
        }
        t.v[k - N] = (uint32_t)acc & kLimbMask;
        acc >>= kLimbBits;
    }
    t.v[N - 1] = (uint32_t)acc;  // the value is below 2^(29N): fits the top limb
    return t;
}

#ifndef KZ_MUL_ILP
#define KZ_MUL_ILP 1
#endif

template <class F>
KZ_DEV Fe<F> fe_mul(const Fe<F>& a, const Fe<F>& b) {
    return fe_mul_k<F, KZ_MUL_ILP>(a, b);
}

// Montgomery square a^2 2^(-29N) mod p: the product scan of fe_mul_k with
// each cross product a_i a_j (i < j) taken once against the doubled limb
// 2 a_i (< 2^30) -- (N^2 + N) / 2 limb products for the a*a part instead of
// N^2 (105 instead of 196 for BLS12-381 Fp), with no extra work per column.
// A column holds at most N/2 cross products below 2^59, one square below 2^58
// and N reduction products below 2^58: below 2^63.
template <class F, bool Canon = true>
KZ_DEV Fe<F> fe_sqr(const Fe<F>& a) {
    constexpr int N = F::N;
    uint32_t m[N], d[N];
    Fe<F> t;
#pragma unroll
    for (int i = 0; i < N; ++i) d[i] = a.v[i] << 1;
    uint64_t acc = 0;
#pragma unroll
    for (int k = 0; k < N; ++k) {
#pragma unroll
        for (int i = 0; 2 * i < k; ++i) acc += (uint64_t)d[i] * a.v[k - i];
        if ((k & 1) == 0) acc += (uint64_t)a.v[k / 2] * a.v[k / 2];
#pragma unroll
        for (int i = 0; i < k; ++i) acc += (uint64_t)m[i] * F::P[k - i];
        m[k] = ((uint32_t)acc * F::PINV) & kLimbMask;
        acc += (uint64_t)m[k] * F::P[0];
        acc >>= kLimbBits;
    }
#pragma unroll
    for (int k = N; k < 2 * N - 1; ++k) {
#pragma unroll
        for (int i = k - N + 1; 2 * i < k; ++i) acc += (uint64_t)d[i] * a.v[k - i];
        if ((k & 1) == 0) acc += (uint64_t)a.v[k / 2] * a.v[k / 2];
#pragma unroll
        for (int i = k - N + 1; i < N; ++i) acc += (uint64_t)m[i] * F::P[k - i];
        t.v[k - N] = (uint32_t)acc & kLimbMask;
        acc >>= kLimbBits;
    }
    t.v[N - 1] = (uint32_t)acc;  // < 2p < 2^(29N): fits the top limb
    if (Canon) fe_reduce_once<F>(t);
    return t;
}

// ------------------------------------------------------- lazy reduction
// Chains of point operations that never compare or store their values skip
// the products' final subtraction (fe_mul_k<F, K, false>, fe_sqr<F, false>).
// The Montgomery product maps inputs a, b with a b < R p to outputs below
// p + a b / R < 2p; R / p is 2^25.3 for BLS12-381 Fp and 2^7.4 (about 169)
// for BN254 Fp, so operands up to 10p (10 * 10 < 169) still give products
// below 2p.  Values below 10p fit N normalised limbs.  Differences then need
// no conditional correction: a + k p - b is computed in one pass with signed
// carries, the caller choosing k p above b's bound.
template <class F, int K>
constexpr uint32_t kp_limb(int i) {
    uint64_t c = 0, t = 0;
    for (int j = 0; j <= i; ++j) {
        t = (uint64_t)K * F::P[j] + c;
        c = t >> kLimbBits;
    }
    return (uint32_t)t & kLimbMask;
}

// a + K p - b for a + K p - b in (0, 2^(29N)) (limbs normalised)
template <class F, int K>
KZ_DEV Fe<F> fe_sub_kp(const Fe<F>& a, const Fe<F>& b) {
    Fe<F> d;
    int32_t c = 0;
#pragma unroll
    for (int i = 0; i < F::N; ++i) {
        const int32_t t = (int32_t)(a.v[i] + kp_limb<F, K>(i)) - (int32_t)b.v[i] + c;
        c = t >> kLimbBits;
        d.v[i] = (uint32_t)t & kLimbMask;
    }
    return d;
}

// r + 6p - s - 2q in one pass, for r, s, q < 2p: a value in (0, 8p)
template <class F>
KZ_DEV Fe<F> fe_r6p_s_2q(const Fe<F>& r, const Fe<F>& s, const Fe<F>& q) {
    Fe<F> d;
    int32_t c = 0;
#pragma unroll
    for (int i = 0; i < F::N; ++i) {
        const int32_t t = (int32_t)(r.v[i] + kp_limb<F, 6>(i)) - (int32_t)s.v[i] - (int32_t)(q.v[i] << 1) + c;
        c = t >> kLimbBits;
        d.v[i] = (uint32_t)t & kLimbMask;
    }
    return d;
}

// a + b with normalised limbs and no reduction (a + b < 2^(29N)).
template <class F>
KZ_DEV Fe<F> fe_add_nored(const Fe<F>& a, const Fe<F>& b) {
    Fe<F> s;
    uint32_t carry = 0;
#pragma unroll
    for (int i = 0; i < F::N; ++i) {
        const uint32_t t = a.v[i] + b.v[i] + carry;
        carry = t >> kLimbBits;
        s.v[i] = t & kLimbMask;
    }
    return s;
}

// a + K p - b - c in one pass (caller: the value lies in (0, 2^(29N)))
template <class F, int K>
KZ_DEV Fe<F> fe_kp_sub2(const Fe<F>& a, const Fe<F>& b, const Fe<F>& c) {
    Fe<F> d;
    int32_t cy = 0;
#pragma unroll
    for (int i = 0; i < F::N; ++i) {
        const int32_t t = (int32_t)(a.v[i] + kp_limb<F, K>(i)) - (int32_t)b.v[i] - (int32_t)c.v[i] + cy;
        cy = t >> kLimbBits;
        d.v[i] = (uint32_t)t & kLimbMask;
    }
    return d;
}

// a + K p - 2b in one pass (caller: the value lies in (0, 2^(29N)))
template <class F, int K>
KZ_DEV Fe<F> fe_kp_2b(const Fe<F>& a, const Fe<F>& b) {
    Fe<F> d;
    int32_t c = 0;
#pragma unroll
    for (int i = 0; i < F::N; ++i) {
        const int32_t t = (int32_t)(a.v[i] + kp_limb<F, K>(i)) - (int32_t)(b.v[i] << 1) + c;
        c = t >> kLimbBits;
        d.v[i] = (uint32_t)t & kLimbMask;
    }
    return d;
}

// m a for m <= 7 (limb products stay below 2^32) and m a < 2^(29N), no reduction
template <class F, int Mul>
KZ_DEV Fe<F> fe_small_mul(const Fe<F>& a) {
    Fe<F> d;
    uint32_t c = 0;
#pragma unroll
    for (int i = 0; i < F::N; ++i) {
        const uint32_t t = a.v[i] * Mul + c;
        c = t >> kLimbBits;
        d.v[i] = t & kLimbMask;
    }
    return d;
}

// a <- a - K p when a >= K p
template <class F, int K>
KZ_DEV void fe_csub_kp(Fe<F>& a) {
    uint32_t d[F::N];
    uint32_t borrow = 0;
#pragma unroll
    for (int i = 0; i < F::N; ++i) {
        const uint32_t t = a.v[i] - kp_limb<F, K>(i) - borrow;
        borrow = t >> 31;
        d[i] = t & kLimbMask;
    }
#pragma unroll
    for (int i = 0; i < F::N; ++i) a.v[i] = borrow ? a.v[i] : d[i];
}

// A value below 2p is zero mod p exactly when it is 0 or p.
template <class F>
KZ_DEV bool fe_is_zero_lazy(const Fe<F>& a) {
    uint32_t z = 0, e = 0;
#pragma unroll
    for (int i = 0; i < F::N; ++i) {
        z |= a.v[i];
        e |= a.v[i] ^ F::P[i];
    }
    return z == 0 || e == 0;
}

template <class F>
KZ_DEV Fe<F> fe_to_mont(const Fe<F>& a) {
    return fe_mul<F>(a, fe_const<F>(F::R2));
}

template <class F>
KZ_DEV Fe<F> fe_from_mont(const Fe<F>& a) {
    Fe<F> one = fe_zero<F>();
    one.v[0] = 1;
    return fe_mul<F>(a, one);
}

// ------------------------------------------------------ integer conversions
// Little-endian 32-bit words (nw words, value < 2^(29N)) to radix-2^29 limbs.
template <class F>
KZ_DEV Fe<F> fe_from_words(const uint32_t* w, int nw) {
    Fe<F> r;
#pragma unroll
    for (int i = 0; i < F::N; ++i) {
        const int bit = kLimbBits * i;
        const int wi = bit >> 5, sh = bit & 31;
        uint32_t v = (wi < nw) ? (w[wi] >> sh) : 0u;
        if (sh > 32 - kLimbBits && wi + 1 < nw) v |= w[wi + 1] << (32 - sh);
        r.v[i] = v & kLimbMask;
    }
    return r;
}

// Radix-2^29 limbs to nw little-endian 32-bit words.
template <class F>
KZ_DEV void fe_to_words(const Fe<F>& a, uint32_t* w, int nw) {
    for (int k = 0; k < nw; ++k) w[k] = 0;
#pragma unroll
    for (int i = 0; i < F::N; ++i) {
        const int bit = kLimbBits * i;
        const int wi = bit >> 5, sh = bit & 31;
        if (wi < nw) w[wi] |= a.v[i] << sh;
        if (sh > 32 - kLimbBits && wi + 1 < nw) w[wi + 1] |= a.v[i] >> (32 - sh);
    }
}

// Big-endian bytes to limbs (value < 2^(29N)).
template <class F>
KZ_DEV Fe<F> fe_from_be_bytes(const uint8_t* b, int nbytes) {
    uint32_t w[16];
    const int nw = (nbytes + 3) / 4;
    for (int k = 0; k < nw; ++k) w[k] = 0;
    for (int k = 0; k < nbytes; ++k) {
        const int bit = 8 * (nbytes - 1 - k);
        w[bit >> 5] |= (uint32_t)b[k] << (bit & 31);
    }
    return fe_from_words<F>(w, nw);
}

template <class F>
KZ_DEV void fe_to_be_bytes(const Fe<F>& a, uint8_t* b, int nbytes) {
    uint32_t w[16];
    const int nw = (nbytes + 3) / 4;
    fe_to_words<F>(a, w, nw);
    for (int k = 0; k < nbytes; ++k) {
        const int bit = 8 * (nbytes - 1 - k);
        b[k] = (uint8_t)(w[bit >> 5] >> (bit & 31));
    }
}

// Integer comparison a > b of normalised values.
template <class F>
KZ_DEV bool fe_gt(const Fe<F>& a, const Fe<F>& b) {
    uint32_t borrow = 0;  // computes b - a; borrow <=> b < a
#pragma unroll
    for (int i = 0; i < F::N; ++i) borrow = (b.v[i] - a.v[i] - borrow) >> 31;
    return borrow != 0;
}

// ------------------------------------------------------------ inversion
// Bernstein-Yang divstep inversion ("Fast constant-time gcd computation and
// modular inversion", 2019), variable time in the number of batches only.
// State: f odd, g, eta = -delta (start f = p, g = A, eta = -1); one divstep:
//   g even:          g <- g/2,                    eta <- eta - 1
//   g odd, eta >= 0: g <- (g + f)/2,              eta <- eta - 1
//   g odd, eta < 0:  (f, g) <- (g, (g - f)/2),    eta <- -eta - 1
// g reaches 0 within (49 d + 80)/17 steps for d-bit inputs, and then f = +-1.
// Steps depend only on the low bits of f and g, so batches of 29 run on the
// low limbs alone and record the 2x2 matrix T with 2^29 (f', g') = T (f, g);
// T is then applied to the full values (exact division by 2^29) and to the
// coefficients (d, e), kept with d A = f, e A = g (mod p) and divided by
// 2^29 mod p by adding the multiple of p that clears their low limb.
// Multi-limb values are signed: limbs 0..N-2 in [0, 2^29), the top limb a
// signed int32.  |T entries| <= 2^29, so every limb sum fits an int64.
template <class F>
KZ_DEV void dv_apply(int32_t u, int32_t v, int32_t q, int32_t r, int32_t* f, int32_t* g) {
    int64_t cf = (int64_t)u * f[0] + (int64_t)v * g[0];
    int64_t cg = (int64_t)q * f[0] + (int64_t)r * g[0];
    cf >>= kLimbBits;  // low limbs are zero: exact division
    cg >>= kLimbBits;
#pragma unroll
    for (int i = 1; i < F::N; ++i) {
        cf += (int64_t)u * f[i] + (int64_t)v * g[i];
        cg += (int64_t)q * f[i] + (int64_t)r * g[i];
        f[i - 1] = (int32_t)(cf & kLimbMask);
        g[i - 1] = (int32_t)(cg & kLimbMask);
        cf >>= kLimbBits;
        cg >>= kLimbBits;
    }
    f[F::N - 1] = (int32_t)cf;
    g[F::N - 1] = (int32_t)cg;
}

// d <- d - p if d >= p (d signed, |d| < 2^(29N - 1))
template <class F>
KZ_DEV void dv_sub_p_if_ge(int32_t* d) {
    int32_t t[F::N];
    int32_t c = 0;
#pragma unroll
    for (int i = 0; i < F::N - 1; ++i) {
        const int32_t x = d[i] - (int32_t)F::P[i] + c;
        t[i] = x & (int32_t)kLimbMask;
        c = x >> kLimbBits;
    }
    t[F::N - 1] = d[F::N - 1] - (int32_t)F::P[F::N - 1] + c;
    const bool ge = t[F::N - 1] >= 0;
#pragma unroll
    for (int i = 0; i < F::N; ++i) d[i] = ge ? t[i] : d[i];
}

// (d, e) <- T (d, e) / 2^29 mod p; d, e in (-p, p) before and after.
template <class F>
KZ_DEV void dv_apply_mod(int32_t u, int32_t v, int32_t q, int32_t r, int32_t* d, int32_t* e) {
    int64_t cd = (int64_t)u * d[0] + (int64_t)v * e[0];
    int64_t ce = (int64_t)q * d[0] + (int64_t)r * e[0];
    // md p clears the low limb: md = -cd p^-1 mod 2^29 (PINV = -p^-1)
    const int64_t md = (int64_t)(((uint32_t)cd * F::PINV) & kLimbMask);
    const int64_t me = (int64_t)(((uint32_t)ce * F::PINV) & kLimbMask);
    cd = (cd + md * (int64_t)F::P[0]) >> kLimbBits;
    ce = (ce + me * (int64_t)F::P[0]) >> kLimbBits;
#pragma unroll
    for (int i = 1; i < F::N; ++i) {
        cd += (int64_t)u * d[i] + (int64_t)v * e[i] + md * (int64_t)F::P[i];
        ce += (int64_t)q * d[i] + (int64_t)r * e[i] + me * (int64_t)F::P[i];
        d[i - 1] = (int32_t)(cd & kLimbMask);
        e[i - 1] = (int32_t)(ce & kLimbMask);
        cd >>= kLimbBits;
        ce >>= kLimbBits;
    }
    d[F::N - 1] = (int32_t)cd;
    e[F::N - 1] = (int32_t)ce;
    // |u d + v e| < 2^29 p and 0 <= md p < 2^29 p: now in (-p, 2p)
    dv_sub_p_if_ge<F>(d);
    dv_sub_p_if_ge<F>(e);
}

// Inverse of a Montgomery-form element (a != 0): a^-1 in Montgomery form.
template <class F>
KZ_DEV Fe<F> fe_inv(const Fe<F>& a) {
    if (fe_is_zero<F>(a)) return a;  // no inverse; callers never ask for it
    constexpr int kBatch = kLimbBits;
    constexpr int kMaxBatches = ((49 * F::BITS + 80) / 17 + kBatch - 1) / kBatch + 1;
    int32_t f[F::N], g[F::N], d[F::N], e[F::N];
#pragma unroll
    for (int i = 0; i < F::N; ++i) {
        f[i] = (int32_t)F::P[i];
        g[i] = (int32_t)a.v[i];
        d[i] = 0;
        e[i] = 0;
    }
    e[0] = 1;
    int32_t eta = -1;
#pragma unroll 1
    for (int b = 0; b < kMaxBatches; ++b) {
        int32_t gz = 0;
#pragma unroll
        for (int i = 0; i < F::N; ++i) gz |= g[i];
        if (gz == 0) break;
        // 29 divsteps on the low limbs, branch-free
        uint32_t fl = (uint32_t)f[0], gl = (uint32_t)g[0];
        int32_t u = 1, v = 0, q = 0, r = 1;
#pragma unroll 1
        for (int i = 0; i < kBatch; ++i) {
            const int32_t neg = eta >> 31;            // eta < 0
            const int32_t odd = -(int32_t)(gl & 1u);  // g odd
            // g odd: g += f (eta >= 0) or g -= f (eta < 0); rows likewise
            gl += ((fl ^ (uint32_t)neg) - (uint32_t)neg) & (uint32_t)odd;
            q += ((u ^ neg) - neg) & odd;
            r += ((v ^ neg) - neg) & odd;
            // swap (g odd, eta < 0): f <- old g = f + (g - f), eta <- -eta - 1
            const int32_t sw = neg & odd;
            eta = (eta ^ sw) - (sw + 1);
            fl += gl & (uint32_t)sw;
            u += q & sw;
            v += r & sw;
            gl >>= 1;
            u *= 2;
            v *= 2;
        }
        dv_apply<F>(u, v, q, r, f, g);
        dv_apply_mod<F>(u, v, q, r, d, e);
    }
    // f = +-1: A^-1 = f d, d in (-p, p)
    const int32_t fneg = f[F::N - 1] >> 31;
    int32_t c = 0;
    Fe<F> res;
#pragma unroll
    for (int i = 0; i < F::N; ++i) {
        const int32_t x = ((d[i] ^ fneg) - fneg) + c;  // limbwise negation, signed carries
        const int32_t lim = i + 1 < F::N ? (x & (int32_t)kLimbMask) : x;
        c = i + 1 < F::N ? (x >> kLimbBits) : 0;
        d[i] = lim;
    }
    // now in (-p, p): add p if negative
    const int32_t dneg = d[F::N - 1] >> 31;
    c = 0;
#pragma unroll
    for (int i = 0; i < F::N; ++i) {
        const int32_t x = d[i] + ((int32_t)F::P[i] & dneg) + c;
        res.v[i] = (uint32_t)(i + 1 < F::N ? (x & (int32_t)kLimbMask) : x);
        c = x >> kLimbBits;
    }
    // res = (a R)^-1 as a plain value; its Montgomery product with R^3 is a^-1 R.
    return fe_mul<F>(res, fe_const<F>(F::R3));
}

}  // namespace kz
