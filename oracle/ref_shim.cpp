// extern "C" entry points onto the reference's header-only field and curve code,
// compiled together with the reference sources into oracle/_ref/libref.so.
// Test infrastructure only: used to cross-check the C oracle and to write the
// golden fixtures under tests/golden/.  Nothing here is product code.
//
// Bound reference functions (all under /root/reference):
//   bls/cpp/bls12_381_fp.hpp   fp_mul / to_mont_fp / from_mont_fp
//   bls/cpp/bls12_381_g1.hpp   g1_add / g1_double / g1_scalar_mul_256 / g1_to_affine
//   bn254/cpp/bn254_fp.hpp     fp_mul / fr_mul / to_mont_* / from_mont_*
//   bn254/cpp/bn254_g1.hpp     g1_add / g1_double / g1_scalar_mul
//   kzg/cpp/kzg_oracle.hpp     fr_mont_mul / fr_to_mont / fr_from_mont (BLS12-381 Fr)
// gpukit_multi_pippenger_cpu itself is exported by multi_pippenger.cpp.
#include <cstdint>
#include <cstring>

#include "bls/cpp/bls12_381_g1.hpp"
#include "bn254/cpp/bn254_g1.hpp"
#include "kzg/cpp/kzg_oracle.hpp"

namespace bls = kinet::crypto::bls12_381;
namespace bn = kinet::crypto::bn254;
namespace kz = kinet::crypto::kzg;

namespace {

bls::G1Jac bls_read(const uint8_t* w) {
    bool any = false;
    for (int i = 0; i < 96; ++i) any |= w[i] != 0;
    if (!any) return bls::g1_jac_zero();
    bls::G1Affine a;
    a.x = bls::to_mont_fp(bls::U384::from_be48(w));
    a.y = bls::to_mont_fp(bls::U384::from_be48(w + 48));
    a.infinity = false;
    return bls::g1_to_jac(a);
}

void bls_write(const bls::G1Jac& p, uint8_t* w) {
    bls::G1Affine a = bls::g1_to_affine(p);
    if (a.infinity) { std::memset(w, 0, 96); return; }
    bls::from_mont_fp(a.x).to_be48(w);
    bls::from_mont_fp(a.y).to_be48(w + 48);
}

bn::G1Jac bn_read(const uint8_t* w) {
    bool any = false;
    for (int i = 0; i < 64; ++i) any |= w[i] != 0;
    if (!any) return bn::g1_jac_zero();
    bn::G1Affine a;
    a.x = bn::to_mont_fp(bn::U256::from_be32(w));
    a.y = bn::to_mont_fp(bn::U256::from_be32(w + 32));
    a.infinity = false;
    return bn::g1_to_jac(a);
}

void bn_write(const bn::G1Jac& p, uint8_t* w) {
    bn::G1Affine a = bn::g1_to_affine(p);
    if (a.infinity) { std::memset(w, 0, 64); return; }
    bn::from_mont_fp(a.x).to_be32(w);
    bn::from_mont_fp(a.y).to_be32(w + 32);
}

template <class U, int N>
U from_le(const uint8_t* b) {
    U r;
    for (int l = 0; l < N; ++l) {
        uint64_t v = 0;
        for (int k = 7; k >= 0; --k) v = (v << 8) | b[8 * l + k];
        r.limbs[l] = v;
    }
    return r;
}

template <class U, int N>
void to_le(const U& a, uint8_t* b) {
    for (int l = 0; l < N; ++l)
        for (int k = 0; k < 8; ++k) b[8 * l + k] = (uint8_t)(a.limbs[l] >> (8 * k));
}

}  // namespace

extern "C" {

// canonical LE in/out; field 0 = BLS Fp, 1 = BLS Fr, 2 = BN Fp, 3 = BN Fr
void ref_field_mul(int field, const uint8_t* a, const uint8_t* b, uint8_t* out) {
    switch (field) {
        case 0: {
            auto x = bls::to_mont_fp(from_le<bls::U384, 6>(a));
            auto y = bls::to_mont_fp(from_le<bls::U384, 6>(b));
            to_le<bls::U384, 6>(bls::from_mont_fp(bls::fp_mul(x, y)), out);
            break;
        }
        case 1: {
            kz::Fr x, y;
            for (int l = 0; l < 4; ++l) {
                x.limbs[l] = from_le<bn::U256, 4>(a).limbs[l];
                y.limbs[l] = from_le<bn::U256, 4>(b).limbs[l];
            }
            kz::Fr r = kz::fr_from_mont(kz::fr_mont_mul(kz::fr_to_mont(x), kz::fr_to_mont(y)));
            bn::U256 u;
            for (int l = 0; l < 4; ++l) u.limbs[l] = r.limbs[l];
            to_le<bn::U256, 4>(u, out);
            break;
        }
        case 2: {
            auto x = bn::to_mont_fp(from_le<bn::U256, 4>(a));
            auto y = bn::to_mont_fp(from_le<bn::U256, 4>(b));
            to_le<bn::U256, 4>(bn::from_mont_fp(bn::fp_mul(x, y)), out);
            break;
        }
        case 3: {
            auto x = bn::to_mont_fr(from_le<bn::U256, 4>(a));
            auto y = bn::to_mont_fr(from_le<bn::U256, 4>(b));
            to_le<bn::U256, 4>(bn::from_mont_fr(bn::fr_mul(x, y)), out);
            break;
        }
    }
}

// curve 1 = BN254, 2 = BLS12-381; wire = BE affine, zero = identity
void ref_g1_add(int curve, const uint8_t* a, const uint8_t* b, uint8_t* out) {
    if (curve == 2) bls_write(bls::g1_add(bls_read(a), bls_read(b)), out);
    else bn_write(bn::g1_add(bn_read(a), bn_read(b)), out);
}

void ref_g1_dbl(int curve, const uint8_t* a, uint8_t* out) {
    if (curve == 2) bls_write(bls::g1_double(bls_read(a)), out);
    else bn_write(bn::g1_double(bn_read(a)), out);
}

// k: 32-byte LE
void ref_g1_mul(int curve, const uint8_t* p, const uint8_t* k, uint8_t* out) {
    if (curve == 2) {
        uint64_t kl[4];
        for (int l = 0; l < 4; ++l) kl[l] = from_le<bn::U256, 4>(k).limbs[l];
        bls::G1Jac j = bls_read(p);
        bls::G1Affine a = bls::g1_to_affine(j);
        bls_write(bls::g1_scalar_mul_256(a, kl), out);
    } else {
        bn::G1Jac j = bn_read(p);
        bn::G1Affine a = bn::g1_to_affine(j);
        bn_write(bn::g1_scalar_mul(a, from_le<bn::U256, 4>(k)), out);
    }
}

void ref_g1_generator(int curve, uint8_t* out) {
    if (curve == 2) {
        bls_write(bls::g1_to_jac(bls::g1_generator()), out);
    } else {
        std::memset(out, 0, 64);
        out[31] = 1;
        out[63] = 2;
    }
}

}  // extern "C"
