// Host build of the library's device arithmetic (crypto_amd/csrc/field.h,
// curve.h), exported for the CPU tests, which compare it with the oracle.
// Test infrastructure only.
#include <cstring>

#include "curve.h"
#include "glv.h"

using namespace kz;

namespace {

template <class F>
Fe<F> from_le(const uint8_t* b, int nbytes) {
    uint32_t w[16] = {0};
    for (int k = 0; k < nbytes; ++k) w[k / 4] |= (uint32_t)b[k] << (8 * (k % 4));
    return fe_from_words<F>(w, (nbytes + 3) / 4);
}

template <class F>
void to_le(const Fe<F>& a, uint8_t* b, int nbytes) {
    uint32_t w[16] = {0};
    fe_to_words<F>(a, w, (nbytes + 3) / 4);
    for (int k = 0; k < nbytes; ++k) b[k] = (uint8_t)(w[k / 4] >> (8 * (k % 4)));
}

// op: 0 add 1 sub 2 mul 3 inv 4 square of a; canonical LE in/out
template <class F>
void field_op(int op, const uint8_t* a, const uint8_t* b, uint8_t* out, int nbytes) {
    Fe<F> x = fe_to_mont<F>(from_le<F>(a, nbytes)), y = fe_to_mont<F>(from_le<F>(b, nbytes)), r;
    switch (op) {
        case 0: r = fe_add<F>(x, y); break;
        case 1: r = fe_sub<F>(x, y); break;
        case 2: r = fe_mul<F>(x, y); break;
        case 4: r = fe_sqr<F>(x); break;
        default: r = fe_inv<F>(x); break;
    }
    to_le<F>(fe_from_mont<F>(r), out, nbytes);
}

template <class C>
Xyzz<typename C::Fp> read_wire(const uint8_t* w) {
    using F = typename C::Fp;
    bool any = false;
    for (int i = 0; i < 2 * C::FB; ++i) any |= w[i] != 0;
    if (!any) return xyzz_inf<F>();
    Affine<F> a;
    a.x = fe_to_mont<F>(fe_from_be_bytes<F>(w, C::FB));
    a.y = fe_to_mont<F>(fe_from_be_bytes<F>(w + C::FB, C::FB));
    return xyzz_from_affine<F>(a);
}

template <class C>
void write_wire(const Xyzz<typename C::Fp>& p, uint8_t* w) {
    using F = typename C::Fp;
    Affine<F> a;
    bool inf;
    xyzz_to_affine<F>(p, a, inf);
    if (inf) { std::memset(w, 0, 2 * C::FB); return; }
    fe_to_be_bytes<F>(fe_from_mont<F>(a.x), w, C::FB);
    fe_to_be_bytes<F>(fe_from_mont<F>(a.y), w + C::FB, C::FB);
}

// op: 0 general add, 1 mixed add (b affine), 2 double a
template <class C>
void point_op(int op, const uint8_t* a, const uint8_t* b, uint8_t* out) {
    using F = typename C::Fp;
    Xyzz<F> p = read_wire<C>(a), r;
    if (op == 0) {
        r = xyzz_add<F>(p, read_wire<C>(b));
    } else if (op == 1) {
        Xyzz<F> q = read_wire<C>(b);
        Affine<F> qa;
        qa.x = q.X;
        qa.y = q.Y;
        r = p;
        xyzz_madd<F>(r, qa);
    } else if (op == 3) {  // p + 3 q through three mixed additions of the accumulation loop
        Xyzz<F> q = read_wire<C>(b);
        Affine<F> qa;
        qa.x = q.X;
        qa.y = q.Y;
        r = p;
        for (int k = 0; k < 3; ++k) xyzz_madd_lazy<F>(r, qa);
        r = xyzz_canon_acc<F>(r);
    } else if (op == 5) {  // p + 20 q: the loop's bounds in their steady state
        Xyzz<F> q = read_wire<C>(b);
        Affine<F> qa;
        qa.x = q.X;
        qa.y = q.Y;
        r = p;
        for (int k = 0; k < 20; ++k) xyzz_madd_lazy<F>(r, qa);
        r = xyzz_canon_acc<F>(r);
    } else {
        r = xyzz_dbl<F>(p);
    }
    write_wire<C>(r, out);
}

}  // namespace

extern "C" {

// field: 0 BLS Fp, 1 BLS Fr, 2 BN Fp, 3 BN Fr, 4 secp256k1 Fp
void hc_field_op(int field, int op, const uint8_t* a, const uint8_t* b, uint8_t* out) {
    switch (field) {
        case 0: field_op<BlsFp>(op, a, b, out, 48); break;
        case 1: field_op<BlsFr>(op, a, b, out, 32); break;
        case 2: field_op<BnFp>(op, a, b, out, 32); break;
        case 4: field_op<SecpFp>(op, a, b, out, 32); break;
        default: field_op<BnFr>(op, a, b, out, 32); break;
    }
}

// BN254 GLV split of a 32-byte LE scalar: out = |k1| (16 B LE) || |k2| (16 B)
// || sign of k1 || sign of k2
void hc_glv_split_bn(const uint8_t* s32, uint8_t* out) {
    uint32_t s[8] = {0}, k1[4], k2[4];
    for (int k = 0; k < 32; ++k) s[k / 4] |= (uint32_t)s32[k] << (8 * (k % 4));
    bool n1, n2;
    glvbn::split(s, k1, k2, n1, n2);
    for (int k = 0; k < 16; ++k) {
        out[k] = (uint8_t)(k1[k / 4] >> (8 * (k % 4)));
        out[16 + k] = (uint8_t)(k2[k / 4] >> (8 * (k % 4)));
    }
    out[32] = n1 ? 1 : 0;
    out[33] = n2 ? 1 : 0;
}

void hc_glv_split_bls(const uint8_t* s32, uint8_t* out) {
    uint32_t s[8] = {0}, k1[4], k2[4];
    for (int k = 0; k < 32; ++k) s[k / 4] |= (uint32_t)s32[k] << (8 * (k % 4));
    bool n1, n2;
    glvbls::split(s, k1, k2, n1, n2);
    for (int k = 0; k < 16; ++k) {
        out[k] = (uint8_t)(k1[k / 4] >> (8 * (k % 4)));
        out[16 + k] = (uint8_t)(k2[k / 4] >> (8 * (k % 4)));
    }
    out[32] = n1 ? 1 : 0;
    out[33] = n2 ? 1 : 0;
}

void hc_point_op(int curve, int op, const uint8_t* a, const uint8_t* b, uint8_t* out) {
    if (curve == 2) point_op<CurveBLS12381>(op, a, b, out);
    else if (curve == 1) point_op<CurveBN254>(op, a, b, out);
    else point_op<CurveSecp256k1>(op, a, b, out);
}

}  // extern "C"
