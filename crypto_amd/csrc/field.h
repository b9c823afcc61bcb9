// Device finite-field arithmetic for gfx950: Montgomery form over 32-bit limbs.
//
// Four fields, all with the top bit of the modulus free (p < 2^(32N-1) and the
// top limb < 0x7FFFFFFE), which lets the CIOS product run without the extra
// carry word: the interleaved "no-carry" CIOS (Koc-Acar-Kaliski CIOS with the
// spare-bit shortcut).  Products are 32x32->64 multiply-adds, which hipcc lowers
// to v_mad_u64_u32 chains; additions use v_add_co/v_addc_co carry chains.
//
// Values are kept fully reduced (< p) between operations, so every field
// element has a unique representation and equality is limb equality.
//
// Field constants restate the reference's moduli:
//   BLS12-381 Fp  bls/cpp/bls12_381_fp.hpp:83-98
//   BLS12-381 Fr  kzg/cpp/kzg_oracle.hpp:66-90
//   BN254 Fp, Fr  bn254/cpp/bn254_fp.hpp:75-97
// split into 32-bit limbs (R = 2^(32N) is the same value as the reference's
// R = 2^(64n) for n = N/2 u64 limbs, so Montgomery forms coincide).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define KZ_DEV __device__ __forceinline__
#define KZ_HD __host__ __device__ __forceinline__

namespace kz {

struct BlsFp {
    static constexpr int N = 12;
    static constexpr uint32_t PINV = 0xFFFCFFFDu;
    static constexpr uint32_t P[12] = {0xFFFFAAABu, 0xB9FEFFFFu, 0xB153FFFFu, 0x1EABFFFEu,
                                       0xF6B0F624u, 0x6730D2A0u, 0xF38512BFu, 0x64774B84u,
                                       0x434BACD7u, 0x4B1BA7B6u, 0x397FE69Au, 0x1A0111EAu};
    static constexpr uint32_t ONE[12] = {0x0002FFFDu, 0x76090000u, 0xC40C0002u, 0xEBF4000Bu,
                                         0x53C758BAu, 0x5F489857u, 0x70525745u, 0x77CE5853u,
                                         0xA256EC6Du, 0x5C071A97u, 0xFA80E493u, 0x15F65EC3u};
    static constexpr uint32_t R2[12] = {0x1C341746u, 0xF4DF1F34u, 0x09D104F1u, 0x0A76E6A6u,
                                        0x4C95B6D5u, 0x8DE5476Cu, 0x939D83C0u, 0x67EB88A9u,
                                        0xB519952Du, 0x9A793E85u, 0x92CAE3AAu, 0x11988FE5u};
    static constexpr uint32_t R3[12] = {0xD94CA1E0u, 0xED48AC6Bu, 0x03A7ADF8u, 0x315F831Eu,
                                        0x615E29DDu, 0x9A53352Au, 0x921E1761u, 0x34C04E5Eu,
                                        0x65724728u, 0x2512D435u, 0x91755D4Du, 0x0AA63460u};
};

struct BlsFr {
    static constexpr int N = 8;
    static constexpr uint32_t PINV = 0xFFFFFFFFu;
    static constexpr uint32_t P[8] = {0x00000001u, 0xFFFFFFFFu, 0xFFFE5BFEu, 0x53BDA402u,
                                      0x09A1D805u, 0x3339D808u, 0x299D7D48u, 0x73EDA753u};
    static constexpr uint32_t ONE[8] = {0xFFFFFFFEu, 0x00000001u, 0x00034802u, 0x5884B7FAu,
                                        0xECBC4FF5u, 0x998C4FEFu, 0xACC5056Fu, 0x1824B159u};
    static constexpr uint32_t R2[8] = {0xF3F29C6Du, 0xC999E990u, 0x87925C23u, 0x2B6CEDCBu,
                                       0x7254398Fu, 0x05D31496u, 0x9F59FF11u, 0x0748D9D9u};
    static constexpr uint32_t R3[8] = {0x439B73AFu, 0xC62C1807u, 0x8CF06990u, 0x1B3E0D18u,
                                       0xC7B5F418u, 0x73D13C71u, 0xC8DB33E9u, 0x6E2A5BB9u};
};

struct BnFp {
    static constexpr int N = 8;
    static constexpr uint32_t PINV = 0xE4866389u;
    static constexpr uint32_t P[8] = {0xD87CFD47u, 0x3C208C16u, 0x6871CA8Du, 0x97816A91u,
                                      0x8181585Du, 0xB85045B6u, 0xE131A029u, 0x30644E72u};
    static constexpr uint32_t ONE[8] = {0xC58F0D9Du, 0xD35D438Du, 0xF5C70B3Du, 0x0A78EB28u,
                                        0x7879462Cu, 0x666EA36Fu, 0x9A07DF2Fu, 0x0E0A77C1u};
    static constexpr uint32_t R2[8] = {0x538AFA89u, 0xF32CFC5Bu, 0xD44501FBu, 0xB5E71911u,
                                       0x0A417FF6u, 0x47AB1EFFu, 0xCAB8351Fu, 0x06D89F71u};
    static constexpr uint32_t R3[8] = {0xDA1530DFu, 0xB1CD6DAFu, 0xA7283DB6u, 0x62F210E6u,
                                       0x0ADA0AFBu, 0xEF7F0B0Cu, 0x2D592544u, 0x20FD6E90u};
};

struct BnFr {
    static constexpr int N = 8;
    static constexpr uint32_t PINV = 0xEFFFFFFFu;
    static constexpr uint32_t P[8] = {0xF0000001u, 0x43E1F593u, 0x79B97091u, 0x2833E848u,
                                      0x8181585Du, 0xB85045B6u, 0xE131A029u, 0x30644E72u};
    static constexpr uint32_t ONE[8] = {0x4FFFFFFBu, 0xAC96341Cu, 0x9F60CD29u, 0x36FC7695u,
                                        0x7879462Eu, 0x666EA36Fu, 0x9A07DF2Fu, 0x0E0A77C1u};
    static constexpr uint32_t R2[8] = {0xAE216DA7u, 0x1BB8E645u, 0xE35C59E3u, 0x53FE3AB1u,
                                       0x53BB8085u, 0x8C49833Du, 0x7F4E44A5u, 0x0216D0B1u};
    static constexpr uint32_t R3[8] = {0xB4BF0040u, 0x5E94D8E1u, 0x1CFBB6B8u, 0x2A489CBEu,
                                       0xA19FCFEDu, 0x893CC664u, 0x7FCC657Cu, 0x0CF8594Bu};
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
KZ_DEV Fe<F> fe_one() {
    Fe<F> r;
#pragma unroll
    for (int i = 0; i < F::N; ++i) r.v[i] = F::ONE[i];
    return r;
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

// r = a - p if a >= p else a (a < 2p).
template <class F>
KZ_DEV void fe_reduce_once(Fe<F>& a) {
    uint32_t d[F::N];
    uint32_t borrow = 0;
#pragma unroll
    for (int i = 0; i < F::N; ++i) {
        uint64_t s = (uint64_t)a.v[i] - F::P[i] - borrow;
        d[i] = (uint32_t)s;
        borrow = (uint32_t)(s >> 32) & 1u;
    }
#pragma unroll
    for (int i = 0; i < F::N; ++i) a.v[i] = borrow ? a.v[i] : d[i];
}

template <class F>
KZ_DEV Fe<F> fe_add(const Fe<F>& a, const Fe<F>& b) {
    Fe<F> s;
    uint32_t carry = 0;
#pragma unroll
    for (int i = 0; i < F::N; ++i) {
        uint64_t t = (uint64_t)a.v[i] + b.v[i] + carry;
        s.v[i] = (uint32_t)t;
        carry = (uint32_t)(t >> 32);
    }
    // a + b < 2p < 2^(32N): no carry out of the top limb.
    fe_reduce_once<F>(s);
    return s;
}

template <class F>
KZ_DEV Fe<F> fe_sub(const Fe<F>& a, const Fe<F>& b) {
    Fe<F> d, e;
    uint32_t borrow = 0;
#pragma unroll
    for (int i = 0; i < F::N; ++i) {
        uint64_t t = (uint64_t)a.v[i] - b.v[i] - borrow;
        d.v[i] = (uint32_t)t;
        borrow = (uint32_t)(t >> 32) & 1u;
    }
    uint32_t carry = 0;
#pragma unroll
    for (int i = 0; i < F::N; ++i) {
        uint64_t t = (uint64_t)d.v[i] + F::P[i] + carry;
        e.v[i] = (uint32_t)t;
        carry = (uint32_t)(t >> 32);
    }
#pragma unroll
    for (int i = 0; i < F::N; ++i) d.v[i] = borrow ? e.v[i] : d.v[i];
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

// Montgomery product a*b*2^(-32N) mod p, no-carry CIOS.  a, b < p -> result < p.
template <class F>
KZ_DEV Fe<F> fe_mul(const Fe<F>& a, const Fe<F>& b) {
    constexpr int N = F::N;
    uint32_t t[N];
#pragma unroll
    for (int j = 0; j < N; ++j) t[j] = 0;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const uint32_t bi = b.v[i];
        uint64_t A = (uint64_t)a.v[0] * bi + t[0];
        t[0] = (uint32_t)A;
        const uint32_t m = t[0] * F::PINV;
        uint64_t C = (uint64_t)m * F::P[0] + t[0];
#pragma unroll
        for (int j = 1; j < N; ++j) {
            A = (uint64_t)a.v[j] * bi + t[j] + (A >> 32);
            t[j] = (uint32_t)A;
            C = (uint64_t)m * F::P[j] + t[j] + (C >> 32);
            t[j - 1] = (uint32_t)C;
        }
        t[N - 1] = (uint32_t)(C >> 32) + (uint32_t)(A >> 32);
    }
    Fe<F> r;
#pragma unroll
    for (int j = 0; j < N; ++j) r.v[j] = t[j];
    fe_reduce_once<F>(r);
    return r;
}

template <class F>
KZ_DEV Fe<F> fe_sqr(const Fe<F>& a) {
    return fe_mul<F>(a, a);
}

template <class F>
KZ_DEV Fe<F> fe_to_mont(const Fe<F>& a) {
    Fe<F> r2;
#pragma unroll
    for (int i = 0; i < F::N; ++i) r2.v[i] = F::R2[i];
    return fe_mul<F>(a, r2);
}

template <class F>
KZ_DEV Fe<F> fe_from_mont(const Fe<F>& a) {
    Fe<F> one = fe_zero<F>();
    one.v[0] = 1;
    return fe_mul<F>(a, one);
}

// Is a (canonical) < p?
template <class F>
KZ_DEV bool fe_is_canonical(const Fe<F>& a) {
    uint32_t borrow = 0;
#pragma unroll
    for (int i = 0; i < F::N; ++i) {
        uint64_t s = (uint64_t)a.v[i] - F::P[i] - borrow;
        borrow = (uint32_t)(s >> 32) & 1u;
    }
    return borrow != 0;
}

// ------------------------------------------------------------ inversion
// Variable-time binary extended Euclid on the plain integer value of a
// (HAC 14.61 specialised to an odd prime modulus).  Used only on a handful of
// elements (final normalisation), where its ~2*bits short iterations beat a
// ~1.5*bits-multiplication Fermat chain by a wide margin in latency.
template <class F>
KZ_DEV void mp_shr1(uint32_t* a, uint32_t top) {
#pragma unroll
    for (int i = 0; i < F::N - 1; ++i) a[i] = (a[i] >> 1) | (a[i + 1] << 31);
    a[F::N - 1] = (a[F::N - 1] >> 1) | (top << 31);
}

template <class F>
KZ_DEV bool mp_is_one(const uint32_t* a) {
    uint32_t acc = a[0] ^ 1u;
#pragma unroll
    for (int i = 1; i < F::N; ++i) acc |= a[i];
    return acc == 0;
}

template <class F>
KZ_DEV bool mp_geq(const uint32_t* a, const uint32_t* b) {
    uint32_t borrow = 0;
#pragma unroll
    for (int i = 0; i < F::N; ++i) {
        uint64_t s = (uint64_t)a[i] - b[i] - borrow;
        borrow = (uint32_t)(s >> 32) & 1u;
    }
    return borrow == 0;
}

template <class F>
KZ_DEV void mp_sub_inplace(uint32_t* a, const uint32_t* b) {
    uint32_t borrow = 0;
#pragma unroll
    for (int i = 0; i < F::N; ++i) {
        uint64_t s = (uint64_t)a[i] - b[i] - borrow;
        a[i] = (uint32_t)s;
        borrow = (uint32_t)(s >> 32) & 1u;
    }
}

// x <- x/2 mod p for x < p.
template <class F>
KZ_DEV void fe_half_plain(uint32_t* x) {
    if (x[0] & 1u) {
        uint32_t carry = 0;
#pragma unroll
        for (int i = 0; i < F::N; ++i) {
            uint64_t t = (uint64_t)x[i] + F::P[i] + carry;
            x[i] = (uint32_t)t;
            carry = (uint32_t)(t >> 32);
        }
        mp_shr1<F>(x, carry);
    } else {
        mp_shr1<F>(x, 0);
    }
}

// Inverse of a Montgomery-form element (a != 0): returns a^-1 in Montgomery form.
template <class F>
KZ_DEV Fe<F> fe_inv(const Fe<F>& a) {
    if (fe_is_zero<F>(a)) return a;  // no inverse; callers never ask for it
    uint32_t u[F::N], v[F::N], x1[F::N], x2[F::N];
#pragma unroll
    for (int i = 0; i < F::N; ++i) {
        u[i] = a.v[i];
        v[i] = F::P[i];
        x1[i] = 0;
        x2[i] = 0;
    }
    x1[0] = 1;
    while (!mp_is_one<F>(u) && !mp_is_one<F>(v)) {
        while ((u[0] & 1u) == 0) {
            mp_shr1<F>(u, 0);
            fe_half_plain<F>(x1);
        }
        while ((v[0] & 1u) == 0) {
            mp_shr1<F>(v, 0);
            fe_half_plain<F>(x2);
        }
        if (mp_geq<F>(u, v)) {
            mp_sub_inplace<F>(u, v);
            Fe<F> t1, t2;
#pragma unroll
            for (int i = 0; i < F::N; ++i) { t1.v[i] = x1[i]; t2.v[i] = x2[i]; }
            t1 = fe_sub<F>(t1, t2);
#pragma unroll
            for (int i = 0; i < F::N; ++i) x1[i] = t1.v[i];
        } else {
            mp_sub_inplace<F>(v, u);
            Fe<F> t1, t2;
#pragma unroll
            for (int i = 0; i < F::N; ++i) { t1.v[i] = x1[i]; t2.v[i] = x2[i]; }
            t2 = fe_sub<F>(t2, t1);
#pragma unroll
            for (int i = 0; i < F::N; ++i) x2[i] = t2.v[i];
        }
    }
    Fe<F> r;
    const bool uone = mp_is_one<F>(u);
#pragma unroll
    for (int i = 0; i < F::N; ++i) r.v[i] = uone ? x1[i] : x2[i];
    // r = (a R)^-1 as a plain value; times R^3 in Montgomery gives a^-1 R.
    Fe<F> r3;
#pragma unroll
    for (int i = 0; i < F::N; ++i) r3.v[i] = F::R3[i];
    return fe_mul<F>(r, r3);
}

// --------------------------------------------------------------- codecs
// Big-endian bytes (N*4 of them, or nbytes < N*4 with leading limbs zero) to limbs.
template <class F>
KZ_DEV Fe<F> fe_from_be_bytes(const uint8_t* b, int nbytes) {
    Fe<F> r = fe_zero<F>();
    for (int k = 0; k < nbytes; ++k) {
        const int bit = 8 * (nbytes - 1 - k);
        r.v[bit >> 5] |= (uint32_t)b[k] << (bit & 31);
    }
    return r;
}

template <class F>
KZ_DEV void fe_to_be_bytes(const Fe<F>& a, uint8_t* b, int nbytes) {
    for (int k = 0; k < nbytes; ++k) {
        const int bit = 8 * (nbytes - 1 - k);
        b[k] = (uint8_t)(a.v[bit >> 5] >> (bit & 31));
    }
}

}  // namespace kz
