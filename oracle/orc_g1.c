/*
 * Oracle G1 arithmetic for BLS12-381 (y^2 = x^3 + 4) and BN254 (y^2 = x^3 + 3).
 * Test infrastructure only (see oracle.h).
 *
 * Jacobian formulas follow the reference exactly:
 *   g_dbl  <- dbl-2009-l   bls/cpp/bls12_381_g1.hpp:75-110, bn254/cpp/bn254_g1.hpp:66-99
 *   g_add  <- add-2007-bl  bls/cpp/bls12_381_g1.hpp:113-151, bn254/cpp/bn254_g1.hpp:135-170
 *   g_to_affine            bls/cpp/bls12_381_g1.hpp:57-72
 * Wire adapters: gpukit/curve_traits/bls12_381_g1_first_party_traits.h:58-85 and
 * gpukit/cpp/cpu_reference/multi_pippenger.cpp:267-290 (BE x||y, zero = identity).
 * Generators: bls/cpp/bls12_381_g1.hpp:204-224; BN254 (1, 2)
 * (gpukit/test/multi_pippenger_test.cpp:207-211).
 * Compression (no first-party reference code exists; the BLS12-381 form is the
 * ZCash/IETF encoding whose infinity and generator fixtures are in
 * kzg/test/kzg_test.cpp:38-42,71-81).
 */
#include "orc_internal.h"

#include <string.h>

static orc_curve g_curves[3];
static int g_cinit = 0;

static void init_curves(void) {
    if (g_cinit) return;
    memset(g_curves, 0, sizeof g_curves);
    u64 b[6] = {0};
    orc_curve* bn = &g_curves[ORC_CURVE_BN254];
    bn->F = orc_get_field(ORC_FIELD_BN_FP);
    bn->R = orc_get_field(ORC_FIELD_BN_FR);
    bn->fp_bytes = 32;
    b[0] = 3;
    f_to_mont(bn->F, bn->b_mont, b);
    orc_curve* bl = &g_curves[ORC_CURVE_BLS12_381];
    bl->F = orc_get_field(ORC_FIELD_BLS_FP);
    bl->R = orc_get_field(ORC_FIELD_BLS_FR);
    bl->fp_bytes = 48;
    b[0] = 4;
    f_to_mont(bl->F, bl->b_mont, b);
    g_cinit = 1;
}

const orc_curve* orc_get_curve(int curve) {
    init_curves();
    return &g_curves[curve];
}

void g_zero(orc_jac* r) { memset(r, 0, sizeof *r); r->inf = 1; }

void g_dbl(const orc_curve* C, orc_jac* r, const orc_jac* p) {
    const orc_field* F = C->F;
    if (p->inf) { *r = *p; return; }
    if (mp_is_zero(p->Y, F->n)) { g_zero(r); return; }
    u64 A[6], B[6], Cc[6], D[6], E[6], Fq[6], t[6], X3[6], Y3[6], Z3[6];
    f_mul(F, A, p->X, p->X);
    f_mul(F, B, p->Y, p->Y);
    f_mul(F, Cc, B, B);
    f_add(F, t, p->X, B);
    f_mul(F, D, t, t);
    f_sub(F, D, D, A);
    f_sub(F, D, D, Cc);
    f_add(F, D, D, D);
    f_add(F, E, A, A);
    f_add(F, E, E, A);
    f_mul(F, Fq, E, E);
    f_add(F, t, D, D);
    f_sub(F, X3, Fq, t);
    f_sub(F, t, D, X3);
    f_mul(F, Y3, E, t);
    f_add(F, t, Cc, Cc);
    f_add(F, t, t, t);
    f_add(F, t, t, t);
    f_sub(F, Y3, Y3, t);
    f_mul(F, Z3, p->Y, p->Z);
    f_add(F, Z3, Z3, Z3);
    memcpy(r->X, X3, sizeof X3);
    memcpy(r->Y, Y3, sizeof Y3);
    memcpy(r->Z, Z3, sizeof Z3);
    r->inf = 0;
}

void g_add(const orc_curve* C, orc_jac* r, const orc_jac* a, const orc_jac* b) {
    const orc_field* F = C->F;
    if (a->inf) { *r = *b; return; }
    if (b->inf) { *r = *a; return; }
    u64 Z1Z1[6], Z2Z2[6], U1[6], U2[6], S1[6], S2[6], H[6], I[6], J[6], rr[6], V[6], t[6];
    u64 X3[6], Y3[6], Z3[6];
    f_mul(F, Z1Z1, a->Z, a->Z);
    f_mul(F, Z2Z2, b->Z, b->Z);
    f_mul(F, U1, a->X, Z2Z2);
    f_mul(F, U2, b->X, Z1Z1);
    f_mul(F, t, a->Y, b->Z);
    f_mul(F, S1, t, Z2Z2);
    f_mul(F, t, b->Y, a->Z);
    f_mul(F, S2, t, Z1Z1);
    f_sub(F, H, U2, U1);
    if (mp_is_zero(H, F->n)) {
        if (mp_cmp(S1, S2, F->n) == 0) { orc_jac tmp = *a; g_dbl(C, r, &tmp); return; }
        g_zero(r);
        return;
    }
    f_add(F, t, H, H);
    f_mul(F, I, t, t);
    f_mul(F, J, H, I);
    f_sub(F, rr, S2, S1);
    f_add(F, rr, rr, rr);
    f_mul(F, V, U1, I);
    f_mul(F, X3, rr, rr);
    f_sub(F, X3, X3, J);
    f_add(F, t, V, V);
    f_sub(F, X3, X3, t);
    f_sub(F, t, V, X3);
    f_mul(F, Y3, rr, t);
    f_add(F, t, S1, S1);
    f_mul(F, t, t, J);
    f_sub(F, Y3, Y3, t);
    f_add(F, t, a->Z, b->Z);
    f_mul(F, Z3, t, t);
    f_sub(F, Z3, Z3, Z1Z1);
    f_sub(F, Z3, Z3, Z2Z2);
    f_mul(F, Z3, Z3, H);
    memcpy(r->X, X3, sizeof X3);
    memcpy(r->Y, Y3, sizeof Y3);
    memcpy(r->Z, Z3, sizeof Z3);
    r->inf = 0;
}

void g_neg(const orc_curve* C, orc_jac* r, const orc_jac* p) {
    *r = *p;
    if (!p->inf) f_neg(C->F, r->Y, p->Y);
}

/* k * P by left-to-right double-and-add over 256 bits (k as 4 LE u64). */
void g_mul(const orc_curve* C, orc_jac* r, const orc_jac* p, const u64* k4) {
    orc_jac acc;
    g_zero(&acc);
    for (int i = 255; i >= 0; --i) {
        g_dbl(C, &acc, &acc);
        if ((k4[i >> 6] >> (i & 63)) & 1) g_add(C, &acc, &acc, p);
    }
    *r = acc;
}

void g_to_affine(const orc_curve* C, u64* x, u64* y, int* inf, const orc_jac* p) {
    const orc_field* F = C->F;
    if (p->inf || mp_is_zero(p->Z, F->n)) {
        memset(x, 0, 8 * F->n);
        memset(y, 0, 8 * F->n);
        *inf = 1;
        return;
    }
    u64 zi[6], zi2[6], zi3[6];
    f_inv(F, zi, p->Z);
    f_mul(F, zi2, zi, zi);
    f_mul(F, zi3, zi2, zi);
    f_mul(F, x, p->X, zi2);
    f_mul(F, y, p->Y, zi3);
    *inf = 0;
}

void g_read_wire(const orc_curve* C, orc_jac* r, const uint8_t* wire) {
    const int fb = C->fp_bytes;
    int any = 0;
    for (int i = 0; i < 2 * fb; ++i) any |= wire[i];
    if (!any) { g_zero(r); return; }
    u64 x[6] = {0}, y[6] = {0};
    be_to_limbs(x, wire, fb);
    be_to_limbs(y, wire + fb, fb);
    memset(r, 0, sizeof *r);
    f_to_mont(C->F, r->X, x);
    f_to_mont(C->F, r->Y, y);
    memcpy(r->Z, C->F->r1, 8 * C->F->n);
    r->inf = 0;
}

void g_write_wire(const orc_curve* C, uint8_t* wire, const orc_jac* p) {
    const int fb = C->fp_bytes;
    u64 x[6], y[6], xc[6], yc[6];
    int inf;
    g_to_affine(C, x, y, &inf, p);
    if (inf) { memset(wire, 0, 2 * fb); return; }
    f_from_mont(C->F, xc, x);
    f_from_mont(C->F, yc, y);
    limbs_to_be(wire, xc, fb);
    limbs_to_be(wire + fb, yc, fb);
}

void g_generator(const orc_curve* C, orc_jac* r, int curve) {
    static const uint8_t bls_gx[48] = {
        0x17, 0xF1, 0xD3, 0xA7, 0x31, 0x97, 0xD7, 0x94, 0x26, 0x95, 0x63, 0x8C, 0x4F, 0xA9, 0xAC, 0x0F,
        0xC3, 0x68, 0x8C, 0x4F, 0x97, 0x74, 0xB9, 0x05, 0xA1, 0x4E, 0x3A, 0x3F, 0x17, 0x1B, 0xAC, 0x58,
        0x6C, 0x55, 0xE8, 0x3F, 0xF9, 0x7A, 0x1A, 0xEF, 0xFB, 0x3A, 0xF0, 0x0A, 0xDB, 0x22, 0xC6, 0xBB};
    static const uint8_t bls_gy[48] = {
        0x08, 0xB3, 0xF4, 0x81, 0xE3, 0xAA, 0xA0, 0xF1, 0xA0, 0x9E, 0x30, 0xED, 0x74, 0x1D, 0x8A, 0xE4,
        0xFC, 0xF5, 0xE0, 0x95, 0xD5, 0xD0, 0x0A, 0xF6, 0x00, 0xDB, 0x18, 0xCB, 0x2C, 0x04, 0xB3, 0xED,
        0xD0, 0x3C, 0xC7, 0x44, 0xA2, 0x88, 0x8A, 0xE4, 0x0C, 0xAA, 0x23, 0x29, 0x46, 0xC5, 0xE7, 0xE1};
    uint8_t wire[96];
    if (curve == ORC_CURVE_BLS12_381) {
        memcpy(wire, bls_gx, 48);
        memcpy(wire + 48, bls_gy, 48);
    } else {
        memset(wire, 0, 64);
        wire[31] = 1;
        wire[63] = 2;
    }
    g_read_wire(C, r, wire);
}

/* ---- exported ---- */
int orc_point_bytes(int curve) { return 2 * orc_get_curve(curve)->fp_bytes; }
int orc_compressed_bytes(int curve) { return orc_get_curve(curve)->fp_bytes; }

void orc_g1_generator(int curve, uint8_t* out) {
    const orc_curve* C = orc_get_curve(curve);
    orc_jac g;
    g_generator(C, &g, curve);
    g_write_wire(C, out, &g);
}

void orc_g1_mul(int curve, const uint8_t* p_wire, const uint8_t* k_le, uint8_t* out) {
    const orc_curve* C = orc_get_curve(curve);
    orc_jac p, r;
    u64 k[4];
    le_to_limbs(k, k_le, 32);
    g_read_wire(C, &p, p_wire);
    g_mul(C, &r, &p, k);
    g_write_wire(C, out, &r);
}

void orc_g1_mul_gen(int curve, const uint8_t* k_le, uint8_t* out) {
    uint8_t g[96];
    orc_g1_generator(curve, g);
    orc_g1_mul(curve, g, k_le, out);
}

void orc_g1_add(int curve, const uint8_t* a_wire, const uint8_t* b_wire, uint8_t* out) {
    const orc_curve* C = orc_get_curve(curve);
    orc_jac a, b, r;
    g_read_wire(C, &a, a_wire);
    g_read_wire(C, &b, b_wire);
    g_add(C, &r, &a, &b);
    g_write_wire(C, out, &r);
}

void orc_g1_dbl(int curve, const uint8_t* a_wire, uint8_t* out) {
    const orc_curve* C = orc_get_curve(curve);
    orc_jac a, r;
    g_read_wire(C, &a, a_wire);
    g_dbl(C, &r, &a);
    g_write_wire(C, out, &r);
}

int orc_g1_on_curve(int curve, const uint8_t* p_wire) {
    const orc_curve* C = orc_get_curve(curve);
    const orc_field* F = C->F;
    orc_jac p;
    g_read_wire(C, &p, p_wire);
    if (p.inf) return 1;
    u64 lhs[6], rhs[6];
    f_mul(F, lhs, p.Y, p.Y);
    f_mul(F, rhs, p.X, p.X);
    f_mul(F, rhs, rhs, p.X);
    f_add(F, rhs, rhs, C->b_mont);
    return mp_cmp(lhs, rhs, F->n) == 0;
}

/* y is "larger" when y > (p-1)/2, i.e. y > p - y as integers. */
static int y_is_larger(const orc_field* F, const u64* y_canon) {
    u64 ny[6];
    if (mp_is_zero(y_canon, F->n)) return 0;
    mp_sub(ny, F->p, y_canon, F->n);
    return mp_cmp(y_canon, ny, F->n) > 0;
}

void orc_g1_compress(int curve, const uint8_t* p_wire, uint8_t* out) {
    const orc_curve* C = orc_get_curve(curve);
    const orc_field* F = C->F;
    const int fb = C->fp_bytes;
    int any = 0;
    for (int i = 0; i < 2 * fb; ++i) any |= p_wire[i];
    memset(out, 0, fb);
    if (!any) {
        out[0] = (curve == ORC_CURVE_BLS12_381) ? 0xC0 : 0x40;
        return;
    }
    memcpy(out, p_wire, fb);
    u64 y[6] = {0};
    be_to_limbs(y, p_wire + fb, fb);
    int larger = y_is_larger(F, y);
    if (curve == ORC_CURVE_BLS12_381) {
        out[0] |= 0x80;
        if (larger) out[0] |= 0x20;
    } else {
        out[0] |= larger ? 0xC0 : 0x80;
    }
}

int orc_g1_decompress(int curve, const uint8_t* in, uint8_t* out_wire) {
    const orc_curve* C = orc_get_curve(curve);
    const orc_field* F = C->F;
    const int fb = C->fp_bytes;
    uint8_t xb[48];
    int larger;
    memcpy(xb, in, fb);
    /* infinity: the flag byte alone (0xC0 BLS12-381, 0x40 BN254), every
       other bit zero */
    int rest_zero = 1;
    for (int k = 1; k < fb; ++k) rest_zero &= in[k] == 0;
    if (curve == ORC_CURVE_BLS12_381) {
        if (!(in[0] & 0x80)) return -1;
        if (in[0] & 0x40) {
            if (in[0] != 0xC0 || !rest_zero) return -1;
            memset(out_wire, 0, 2 * fb);
            return 0;
        }
        larger = (in[0] & 0x20) != 0;
        xb[0] &= 0x1F;
    } else {
        int flag = in[0] >> 6;
        if (flag == 1) {
            if (in[0] != 0x40 || !rest_zero) return -1;
            memset(out_wire, 0, 2 * fb);
            return 0;
        }
        if (flag < 2) return -1;
        larger = (flag == 3);
        xb[0] &= 0x3F;
    }
    u64 x[6] = {0}, xm[6], rhs[6], ym[6], yc[6];
    be_to_limbs(x, xb, fb);
    if (mp_cmp(x, F->p, F->n) >= 0) return -1;
    f_to_mont(F, xm, x);
    f_mul(F, rhs, xm, xm);
    f_mul(F, rhs, rhs, xm);
    f_add(F, rhs, rhs, C->b_mont);
    if (!f_sqrt(F, ym, rhs)) return -1;
    f_from_mont(F, yc, ym);
    if (y_is_larger(F, yc) != larger) {
        u64 t[6];
        f_neg(F, t, yc);
        memcpy(yc, t, sizeof t);
    }
    memcpy(out_wire, xb, fb);
    limbs_to_be(out_wire + fb, yc, fb);
    return 0;
}
