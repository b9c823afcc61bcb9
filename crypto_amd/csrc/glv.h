// GLV endomorphism split and G1 membership check (glv.hip).
#pragma once

#include "common.h"

namespace kz {

// BN254 GLV split (Gallant-Lambert-Vanstone 2001, with the short lattice basis
// of (r, lambda) from the extended Euclidean algorithm): for any 256-bit s,
// s = k1 + lambda k2 mod r with |k1|, |k2| < 2^127, lambda =
// 0x30644e72e131a029048b6e193fd84104cc37a73fec2bc5e9b8ca0b2d36636f23 (a cube
// root of unity mod r, matched to beta by phi(G) = lambda G).  Basis v1 =
// (A1, -B1), v2 = (A2, B2) with det r; Babai coefficients estimated as
// b1 = floor(k G1 / 2^256), b2 = floor(k G2 / 2^256), G1 = floor(2^256 B2 /
// r), G2 = floor(2^256 B1 / r).  Then (k1, k2) = e1 v1 + e2 v2 with rounding
// errors 0 <= e_i < 1 + k d_i / 2^256 (d_i < 1 the truncation of G_i; for
// k < r the errors stay below 1.016 and 1.082), so 0 <= k1 < 2^126.83 and
// -2^63.13 < k2 < 2^126.92: both magnitudes below 2^127 (also checked
// against big integers over 2e5 random scalars and edge values);
//   k1 = k - b1 A1 - b2 A2,   k2 = b1 B1 - b2 B2   (k = s mod r).
// Writes |k1|, |k2| (4 LE words each) and their signs.  Host and device.
namespace glvbn {
constexpr uint32_t kR[8] = {0xF0000001u, 0x43E1F593u, 0x79B97091u, 0x2833E848u,
                            0x8181585Du, 0xB85045B6u, 0xE131A029u, 0x30644E72u};
constexpr uint32_t kG1[5] = {0x00FF6565u, 0x5398FD03u, 0xA773D2D2u, 0x4CCEF014u, 0x00000002u};
constexpr uint32_t kG2[3] = {0xC7E0B3D7u, 0xD91D232Eu, 0x00000002u};
constexpr uint32_t kA1[4] = {0x7D4F1128u, 0x8211BBEBu, 0xEEB859FCu, 0x6F4D8248u};
constexpr uint32_t kA2[2] = {0x94D213E3u, 0x89D32568u};
constexpr uint32_t kB1[2] = {0x94D213E3u, 0x89D32568u};
constexpr uint32_t kB2[4] = {0x1221250Bu, 0x0BE4E154u, 0xEEB859FDu, 0x6F4D8248u};

// acc (8 words, mod 2^256) += a (NA words) * b (NB words)
template <int NA, int NB>
__host__ __device__ inline void mac8(uint32_t* acc, const uint32_t* a, const uint32_t* b) {
    for (int i = 0; i < NA; ++i) {
        uint64_t carry = 0;
        for (int j = 0; j < NB && i + j < 8; ++j) {
            const uint64_t t = (uint64_t)a[i] * b[j] + acc[i + j] + carry;
            acc[i + j] = (uint32_t)t;
            carry = t >> 32;
        }
        for (int k = i + NB; k < 8 && carry; ++k) {
            const uint64_t t = (uint64_t)acc[k] + carry;
            acc[k] = (uint32_t)t;
            carry = t >> 32;
        }
    }
}

// floor(k g / 2^256) for k (8 words) and g (NG words), result <= 4 words
template <int NG>
__host__ __device__ inline void mul_hi256(const uint32_t* k, const uint32_t* g, uint32_t* out4) {
    uint32_t prod[8 + NG];
    for (int i = 0; i < 8 + NG; ++i) prod[i] = 0;
    for (int i = 0; i < 8; ++i) {
        uint64_t carry = 0;
        for (int j = 0; j < NG; ++j) {
            const uint64_t t = (uint64_t)k[i] * g[j] + prod[i + j] + carry;
            prod[i + j] = (uint32_t)t;
            carry = t >> 32;
        }
        prod[i + NG] = (uint32_t)carry;
    }
    for (int i = 0; i < 4; ++i) out4[i] = 8 + i < 8 + NG ? prod[8 + i] : 0u;
}

// a = -a over 8 words (two's complement)
__host__ __device__ inline void neg8(uint32_t* a) {
    uint64_t c = 1;
    for (int i = 0; i < 8; ++i) {
        const uint64_t t = (uint64_t)(~a[i]) + c;
        a[i] = (uint32_t)t;
        c = t >> 32;
    }
}

__host__ __device__ inline void split(const uint32_t* s_in, uint32_t* k1, uint32_t* k2, bool& neg1, bool& neg2) {
    uint32_t k[8];
    for (int i = 0; i < 8; ++i) k[i] = s_in[i];
    for (int rep = 0; rep < 6; ++rep) {  // k mod r: 2^256 < 6 r
        bool ge = true;
        for (int i = 7; i >= 0; --i)
            if (k[i] != kR[i]) {
                ge = k[i] > kR[i];
                break;
            }
        if (!ge) break;
        uint64_t borrow = 0;
        for (int i = 0; i < 8; ++i) {
            const uint64_t t = (uint64_t)k[i] - kR[i] - borrow;
            k[i] = (uint32_t)t;
            borrow = (t >> 32) & 1u;
        }
    }
    uint32_t b1[4], b2[4];
    mul_hi256<5>(k, kG1, b1);
    mul_hi256<3>(k, kG2, b2);
    // t = b1 A1 + b2 A2;  k1 = k - t
    uint32_t t[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    mac8<4, 4>(t, b1, kA1);
    mac8<2, 2>(t, b2, kA2);
    neg8(t);
    uint32_t x[8];
    {
        uint64_t c = 0;
        for (int i = 0; i < 8; ++i) {
            const uint64_t v = (uint64_t)k[i] + t[i] + c;
            x[i] = (uint32_t)v;
            c = v >> 32;
        }
    }
    neg1 = (x[7] >> 31) != 0;
    if (neg1) neg8(x);
    for (int i = 0; i < 4; ++i) k1[i] = x[i];
    // k2 = b1 B1 - b2 B2
    uint32_t u[8] = {0, 0, 0, 0, 0, 0, 0, 0}, v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    mac8<4, 2>(u, b1, kB1);
    mac8<2, 4>(v, b2, kB2);
    neg8(v);
    uint64_t c = 0;
    for (int i = 0; i < 8; ++i) {
        const uint64_t w = (uint64_t)u[i] + v[i] + c;
        x[i] = (uint32_t)w;
        c = w >> 32;
    }
    neg2 = (x[7] >> 31) != 0;
    if (neg2) neg8(x);
    for (int i = 0; i < 4; ++i) k2[i] = x[i];
}
}  // namespace glvbn

// secp256k1 GLV split (Gallant-Lambert-Vanstone 2001): s = k1 + lambda k2
// mod n for lambda = 0x5363ad4cc05c30e0a5261c028812645a122e22ea20816678df02967c1b23bd72,
// a cube root of unity mod n matched to beta by phi(G) = lambda G.  Basis
// v1 = (A1, -B1), v2 = (A2, B2) of the lattice {(a, b): a + b lambda = 0 mod n}
// from the extended Euclidean algorithm on (n, lambda), det n.  Here the
// basis is longer than BN254's relative to n, so the Babai coefficients are
// rounded at 2^384 precision, b_i = round(k G_i / 2^384) with
// G1 = round(2^384 B2 / n), G2 = round(2^384 B1 / n): the rounding errors stay
// within 1/2 + 2^-128, so |k1| <= (A1 + A2)/2 + 1 < 2^127.35 and
// |k2| <= (B1 + B2)/2 + 1 < 2^127.12 (also checked against big integers over
// 3e5 random scalars and edge values).  Below 2^128 but not 2^127: MSMs over
// this split use 128-bit windows.
//   k1 = k - b1 A1 - b2 A2,   k2 = b1 B1 - b2 B2   (k = s mod n).
// Writes |k1|, |k2| (4 LE words each) and their signs.  Host and device.
namespace glvsecp {
constexpr uint32_t kN[8] = {0xD0364141u, 0xBFD25E8Cu, 0xAF48A03Bu, 0xBAAEDCE6u,
                            0xFFFFFFFEu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu};
constexpr uint32_t kG1[8] = {0x45DBB031u, 0xE893209Au, 0x71E8CA7Fu, 0x3DAA8A14u,
                             0x9284EB15u, 0xE86C90E4u, 0xA7D46BCDu, 0x3086D221u};
constexpr uint32_t kG2[8] = {0x8AC47F71u, 0x1571B4AEu, 0x9DF506C6u, 0x221208ACu,
                             0x0ABFE4C4u, 0x6F547FA9u, 0x010E8828u, 0xE4437ED6u};
constexpr uint32_t kA1[4] = {0x9284EB15u, 0xE86C90E4u, 0xA7D46BCDu, 0x3086D221u};
constexpr uint32_t kA2[5] = {0x9D44CFD8u, 0x57C1108Du, 0xA8E2F3F6u, 0x14CA50F7u, 0x00000001u};
constexpr uint32_t kB1[4] = {0x0ABFE4C3u, 0x6F547FA9u, 0x010E8828u, 0xE4437ED6u};
constexpr uint32_t kB2[4] = {0x9284EB15u, 0xE86C90E4u, 0xA7D46BCDu, 0x3086D221u};

// round(k g / 2^384) for k and g of 8 words; below 2^128 for the G_i above
__host__ __device__ inline void mul_round384(const uint32_t* k, const uint32_t* g, uint32_t* out4) {
    uint32_t prod[16];
    for (int i = 0; i < 16; ++i) prod[i] = 0;
    for (int i = 0; i < 8; ++i) {
        uint64_t carry = 0;
        for (int j = 0; j < 8; ++j) {
            const uint64_t t = (uint64_t)k[i] * g[j] + prod[i + j] + carry;
            prod[i + j] = (uint32_t)t;
            carry = t >> 32;
        }
        prod[i + 8] = (uint32_t)carry;
    }
    uint64_t c = (uint64_t)prod[11] + 0x80000000u;  // + 2^383
    c >>= 32;
    for (int i = 12; i < 16; ++i) {
        c += prod[i];
        out4[i - 12] = (uint32_t)c;
        c >>= 32;
    }
}

__host__ __device__ inline void split(const uint32_t* s_in, uint32_t* k1, uint32_t* k2, bool& neg1, bool& neg2) {
    uint32_t k[8];
    for (int i = 0; i < 8; ++i) k[i] = s_in[i];
    bool ge = true;  // k mod n: 2^256 < 2 n
    for (int i = 7; i >= 0; --i)
        if (k[i] != kN[i]) {
            ge = k[i] > kN[i];
            break;
        }
    if (ge) {
        uint64_t borrow = 0;
        for (int i = 0; i < 8; ++i) {
            const uint64_t t = (uint64_t)k[i] - kN[i] - borrow;
            k[i] = (uint32_t)t;
            borrow = (t >> 32) & 1u;
        }
    }
    uint32_t b1[4], b2[4];
    mul_round384(k, kG1, b1);
    mul_round384(k, kG2, b2);
    // k1 = k - (b1 A1 + b2 A2)  (mod 2^256; the value lies in (-2^128, 2^128))
    uint32_t t[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    glvbn::mac8<4, 4>(t, b1, kA1);
    glvbn::mac8<4, 5>(t, b2, kA2);
    glvbn::neg8(t);
    uint32_t x[8];
    {
        uint64_t c = 0;
        for (int i = 0; i < 8; ++i) {
            const uint64_t v = (uint64_t)k[i] + t[i] + c;
            x[i] = (uint32_t)v;
            c = v >> 32;
        }
    }
    neg1 = (x[7] >> 31) != 0;
    if (neg1) glvbn::neg8(x);
    for (int i = 0; i < 4; ++i) k1[i] = x[i];
    // k2 = b1 B1 - b2 B2
    uint32_t u[8] = {0, 0, 0, 0, 0, 0, 0, 0}, v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    glvbn::mac8<4, 4>(u, b1, kB1);
    glvbn::mac8<4, 4>(v, b2, kB2);
    glvbn::neg8(v);
    uint64_t c = 0;
    for (int i = 0; i < 8; ++i) {
        const uint64_t w = (uint64_t)u[i] + v[i] + c;
        x[i] = (uint32_t)w;
        c = w >> 32;
    }
    neg2 = (x[7] >> 31) != 0;
    if (neg2) glvbn::neg8(x);
    for (int i = 0; i < 4; ++i) k2[i] = x[i];
}
}  // namespace glvsecp

// BLS12-381 GLV split: lambda = 0xac45a4010001a40200000000ffffffff with r =
// lambda^2 + lambda + 1.  For s mod r = k1 + lambda k2 (k1 = s mod lambda,
// k2 = s div lambda <= lambda + 1, both near 2^127.4), two exact moves make
// both parts signed and below 2^127:
//   k2 > (lambda + 1)/2:  (k1, k2) <- (k1 - 1, k2 - lambda - 1)   (s - r)
//   k1 > lambda/2:        (k1, k2) <- (k1 - lambda, k2 + 1)
// so |k1| <= lambda/2 and |k2| <= (lambda + 1)/2 + 1.  With 16-bit windows
// the eight windows of a part are then full and the carry window next to
// them is almost empty, instead of ~26% of the points carrying into it.
// Writes |k1|, |k2| (4 LE words each) and their signs.  Host and device.
namespace glvbls {
constexpr uint32_t kLam[4] = {0xFFFFFFFFu, 0x00000000u, 0x0001A402u, 0xAC45A401u};
constexpr uint32_t kLam1[4] = {0x00000000u, 0x00000001u, 0x0001A402u, 0xAC45A401u};  // lambda + 1
constexpr uint32_t kHalf[4] = {0x7FFFFFFFu, 0x00000000u, 0x8000D201u, 0x5622D200u};  // floor(lambda / 2)
constexpr uint32_t kHalf1[4] = {0x80000000u, 0x00000000u, 0x8000D201u, 0x5622D200u}; // (lambda + 1) / 2
constexpr uint32_t kMu[5] = {0xF6CFEE30u, 0x63F6E522u, 0xE01FAADDu, 0x7C6BECF1u, 0x00000001u};  // 2^256 div lambda
constexpr uint32_t kR[8] = {0x00000001u, 0xFFFFFFFFu, 0xFFFE5BFEu, 0x53BDA402u,
                            0x09A1D805u, 0x3339D808u, 0x299D7D48u, 0x73EDA753u};

// a > b over K words
template <int K>
__host__ __device__ inline bool gt(const uint32_t* a, const uint32_t* b) {
    for (int i = K - 1; i >= 0; --i)
        if (a[i] != b[i]) return a[i] > b[i];
    return false;
}

// a += sign * b over 8 words mod 2^256 (b: K words, zero-extended)
template <int K>
__host__ __device__ inline void addsub8(uint32_t* a, const uint32_t* b, bool subtract) {
    uint64_t c = subtract ? 1u : 0u;
    for (int i = 0; i < 8; ++i) {
        const uint32_t w = i < K ? b[i] : 0u;
        const uint64_t t = (uint64_t)a[i] + (subtract ? ~w : w) + c;
        a[i] = (uint32_t)t;
        c = t >> 32;
    }
}

__host__ __device__ inline void split(const uint32_t* s_in, uint32_t* k1, uint32_t* k2, bool& neg1, bool& neg2) {
    uint32_t s[8];
    for (int i = 0; i < 8; ++i) s[i] = s_in[i];
    for (int rep = 0; rep < 2; ++rep)  // 2^256 < 3 r
        if (!gt<8>(kR, s)) addsub8<8>(s, kR, true);
    // q = floor(s mu / 2^256) <= s div lambda <= q + 2
    uint32_t prod[13];
    for (int i = 0; i < 13; ++i) prod[i] = 0;
    for (int a = 0; a < 8; ++a) {
        uint64_t carry = 0;
        for (int b = 0; b < 5; ++b) {
            const uint64_t t = (uint64_t)s[a] * kMu[b] + prod[a + b] + carry;
            prod[a + b] = (uint32_t)t;
            carry = t >> 32;
        }
        prod[a + 5] = (uint32_t)carry;
    }
    uint32_t q[8] = {prod[8], prod[9], prod[10], prod[11], prod[12], 0, 0, 0};
    // s <- s - q lambda, then at most two more subtractions of lambda
    uint32_t ql[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int a = 0; a < 5; ++a) {
        uint64_t carry = 0;
        for (int b = 0; b < 4 && a + b < 8; ++b) {
            const uint64_t t = (uint64_t)q[a] * kLam[b] + ql[a + b] + carry;
            ql[a + b] = (uint32_t)t;
            carry = t >> 32;
        }
        if (a + 4 < 8) ql[a + 4] = (uint32_t)carry;
    }
    addsub8<8>(s, ql, true);
    const uint32_t one[1] = {1u};
    const uint32_t lam8[8] = {kLam[0], kLam[1], kLam[2], kLam[3], 0, 0, 0, 0};
    for (int rep = 0; rep < 2; ++rep)
        if (!gt<8>(lam8, s)) {
            addsub8<4>(s, kLam, true);
            addsub8<1>(q, one, false);
        }
    // k1 = s (< lambda), k2 = q (<= lambda + 1), as 8-word two's complement
    if (gt<4>(q, kHalf1) || q[4] != 0) {
        addsub8<1>(s, one, true);
        addsub8<4>(q, kLam1, true);
    }
    if (!(s[7] >> 31) && gt<4>(s, kHalf)) {  // s >= 0 here is below 2^128
        addsub8<4>(s, kLam, true);
        addsub8<1>(q, one, false);
    }
    neg1 = (s[7] >> 31) != 0;
    neg2 = (q[7] >> 31) != 0;
    if (neg1) {
        uint32_t t[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        addsub8<8>(t, s, true);
        for (int i = 0; i < 8; ++i) s[i] = t[i];
    }
    if (neg2) {
        uint32_t t[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        addsub8<8>(t, q, true);
        for (int i = 0; i < 8; ++i) q[i] = t[i];
    }
    for (int i = 0; i < 4; ++i) {
        k1[i] = s[i];
        k2[i] = q[i];
    }
}
}  // namespace glvbls

// d_pts2 = [P_i, phi(P_i)] (2n prepared records), d_inf2 likewise, d_scalars2
// = the two parts (16-byte records) of s_i = s1_i + lambda s2_i mod r:
// the signed splits above (magnitudes below 2^127), the signs folded into the
// points (-P, -phi(P)).
// Valid only for points of G1 (all curve points for BN254).
void glv_expand(int curve, const uint32_t* d_pts, const uint8_t* d_inf, const uint32_t* d_scalars, size_t n,
                uint32_t* d_pts2, uint8_t* d_inf2, uint32_t* d_scalars2, hipStream_t s);

// *d_bad |= 1 if a non-identity prepared point is off the curve or, for
// BLS12-381, outside G1 (*d_bad zeroed by the caller).
void check_g1(int curve, const uint32_t* d_pts, const uint8_t* d_inf, size_t n, int* d_bad, hipStream_t s);

}  // namespace kz
