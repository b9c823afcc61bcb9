/*
 * Oracle field arithmetic: Montgomery CIOS over 4 or 6 little-endian u64 limbs.
 * Test infrastructure only (see oracle.h).
 *
 * Follows bls/cpp/bls12_381_fp.hpp:105-239 (add/sub with conditional
 * correction, CIOS mont_mul with a final conditional subtract, Fermat inverse)
 * and bn254/cpp/bn254_fp.hpp (the same CIOS over 4 limbs, fr_mul :239).
 * Moduli are the reference's (bls12_381_fp.hpp:83-85, kzg/cpp/kzg_oracle.hpp:66-71,
 * bn254_fp.hpp:75-81); the Montgomery constants R, R^2, -p^-1 are derived here
 * from p and checked against the reference's tables by the tests.
 */
#include "orc_internal.h"

#include <string.h>

static orc_field g_fields[5];
static int g_init = 0;

/* ---- plain multi-limb helpers ---- */
u64 mp_add(u64* r, const u64* a, const u64* b, int n) {
    u128 c = 0;
    for (int i = 0; i < n; ++i) {
        c += (u128)a[i] + b[i];
        r[i] = (u64)c;
        c >>= 64;
    }
    return (u64)c;
}

u64 mp_sub(u64* r, const u64* a, const u64* b, int n) {
    u64 bw = 0;
    for (int i = 0; i < n; ++i) {
        u128 t = (u128)a[i] - b[i] - bw;
        r[i] = (u64)t;
        bw = (u64)(t >> 64) & 1;
    }
    return bw;
}

int mp_cmp(const u64* a, const u64* b, int n) {
    for (int i = n - 1; i >= 0; --i) {
        if (a[i] < b[i]) return -1;
        if (a[i] > b[i]) return 1;
    }
    return 0;
}

int mp_is_zero(const u64* a, int n) {
    u64 acc = 0;
    for (int i = 0; i < n; ++i) acc |= a[i];
    return acc == 0;
}

/* ---- modular ops (inputs reduced) ---- */
void f_add(const orc_field* F, u64* r, const u64* a, const u64* b) {
    u64 t[6];
    u64 c = mp_add(t, a, b, F->n);
    if (c || mp_cmp(t, F->p, F->n) >= 0) mp_sub(t, t, F->p, F->n);
    memcpy(r, t, 8 * F->n);
}

void f_sub(const orc_field* F, u64* r, const u64* a, const u64* b) {
    u64 t[6];
    u64 bw = mp_sub(t, a, b, F->n);
    if (bw) mp_add(t, t, F->p, F->n);
    memcpy(r, t, 8 * F->n);
}

void f_neg(const orc_field* F, u64* r, const u64* a) {
    if (mp_is_zero(a, F->n)) { memset(r, 0, 8 * F->n); return; }
    mp_sub(r, F->p, a, F->n);
}

/* CIOS Montgomery product a*b*R^-1 mod p (bls12_381_fp.hpp:153-189). */
void f_mul(const orc_field* F, u64* r, const u64* a, const u64* b) {
    const int n = F->n;
    u64 t[8] = {0};
    for (int i = 0; i < n; ++i) {
        u64 carry = 0;
        for (int j = 0; j < n; ++j) {
            u128 pr = (u128)a[j] * b[i] + t[j] + carry;
            t[j] = (u64)pr;
            carry = (u64)(pr >> 64);
        }
        u128 s = (u128)t[n] + carry;
        t[n] = (u64)s;
        t[n + 1] += (u64)(s >> 64);

        u64 m = t[0] * F->pinv;
        carry = 0;
        for (int j = 0; j < n; ++j) {
            u128 pr = (u128)m * F->p[j] + t[j] + carry;
            t[j] = (u64)pr;
            carry = (u64)(pr >> 64);
        }
        s = (u128)t[n] + carry;
        t[n] = (u64)s;
        t[n + 1] += (u64)(s >> 64);
        for (int j = 0; j <= n; ++j) t[j] = t[j + 1];
        t[n + 1] = 0;
    }
    if (t[n] != 0 || mp_cmp(t, F->p, n) >= 0) mp_sub(t, t, F->p, n);
    memcpy(r, t, 8 * n);
}

void f_to_mont(const orc_field* F, u64* r, const u64* a) { f_mul(F, r, a, F->r2); }

void f_from_mont(const orc_field* F, u64* r, const u64* a) {
    u64 one[6] = {1, 0, 0, 0, 0, 0};
    f_mul(F, r, a, one);
}

/* a^e, a in Montgomery form, e plain little-endian with ne limbs. */
void f_pow(const orc_field* F, u64* r, const u64* a, const u64* e, int ne) {
    u64 res[6], base[6];
    memcpy(res, F->r1, 8 * F->n);
    memcpy(base, a, 8 * F->n);
    for (int l = 0; l < ne; ++l) {
        for (int bit = 0; bit < 64; ++bit) {
            if ((e[l] >> bit) & 1) f_mul(F, res, res, base);
            f_mul(F, base, base, base);
        }
    }
    memcpy(r, res, 8 * F->n);
}

/* Fermat inverse a^(p-2) (bls12_381_fp.hpp:234-236). */
void f_inv(const orc_field* F, u64* r, const u64* a) {
    u64 e[6];
    u64 two[6] = {2, 0, 0, 0, 0, 0};
    mp_sub(e, F->p, two, F->n);
    f_pow(F, r, a, e, F->n);
}

/* sqrt for p = 3 mod 4: a^((p+1)/4).  Returns 0 if a is a non-residue. */
int f_sqrt(const orc_field* F, u64* r, const u64* a) {
    u64 e[6];
    u64 one[6] = {1, 0, 0, 0, 0, 0};
    u64 c = mp_add(e, F->p, one, F->n);
    /* shift right by 2, pulling in the carry bit at the top */
    for (int i = 0; i < F->n; ++i) {
        u64 hi = (i + 1 < F->n) ? e[i + 1] : c;
        e[i] = (e[i] >> 2) | (hi << 62);
    }
    u64 cand[6], sq[6];
    f_pow(F, cand, a, e, F->n);
    f_mul(F, sq, cand, cand);
    if (mp_cmp(sq, a, F->n) != 0) return 0;
    memcpy(r, cand, 8 * F->n);
    return 1;
}

/* ---- byte codecs ---- */
void le_to_limbs(u64* r, const uint8_t* b, int nbytes) {
    int n = nbytes / 8;
    for (int i = 0; i < n; ++i) {
        u64 v = 0;
        for (int k = 7; k >= 0; --k) v = (v << 8) | b[8 * i + k];
        r[i] = v;
    }
}

void limbs_to_le(uint8_t* b, const u64* a, int nbytes) {
    int n = nbytes / 8;
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < 8; ++k) b[8 * i + k] = (uint8_t)(a[i] >> (8 * k));
}

void be_to_limbs(u64* r, const uint8_t* b, int nbytes) {
    int n = nbytes / 8;
    for (int l = 0; l < n; ++l) {
        u64 v = 0;
        int base = (n - 1 - l) * 8;
        for (int i = 0; i < 8; ++i) v = (v << 8) | b[base + i];
        r[l] = v;
    }
}

void limbs_to_be(uint8_t* b, const u64* a, int nbytes) {
    int n = nbytes / 8;
    for (int l = 0; l < n; ++l) {
        int base = (n - 1 - l) * 8;
        u64 v = a[l];
        for (int i = 7; i >= 0; --i) { b[base + i] = (uint8_t)v; v >>= 8; }
    }
}

/* ---- field table ---- */
static void derive(orc_field* F) {
    const int n = F->n;
    /* -p^-1 mod 2^64 by Newton iteration on the low limb. */
    u64 inv = 1;
    for (int i = 0; i < 7; ++i) inv *= 2 - F->p[0] * inv;
    F->pinv = (u64)0 - inv;
    /* R mod p = 2^(64n) mod p: double 1 (64n) times with reduction. */
    u64 x[6] = {1, 0, 0, 0, 0, 0};
    for (int i = 0; i < 64 * n; ++i) {
        u64 c = mp_add(x, x, x, n);
        if (c || mp_cmp(x, F->p, n) >= 0) mp_sub(x, x, F->p, n);
    }
    memcpy(F->r1, x, 8 * n);
    for (int i = 0; i < 64 * n; ++i) {
        u64 c = mp_add(x, x, x, n);
        if (c || mp_cmp(x, F->p, n) >= 0) mp_sub(x, x, F->p, n);
    }
    memcpy(F->r2, x, 8 * n);
    F->bits = 0;
    for (int i = 64 * n - 1; i >= 0; --i)
        if ((F->p[i / 64] >> (i % 64)) & 1) { F->bits = i + 1; break; }
}

static void init_fields(void) {
    if (g_init) return;
    /* BLS12-381 base field p (bls/cpp/bls12_381_fp.hpp:83-85). */
    static const u64 bls_p[6] = {0xB9FEFFFFFFFFAAABULL, 0x1EABFFFEB153FFFFULL, 0x6730D2A0F6B0F624ULL,
                                 0x64774B84F38512BFULL, 0x4B1BA7B6434BACD7ULL, 0x1A0111EA397FE69AULL};
    /* BLS12-381 scalar field r (kzg/cpp/kzg_oracle.hpp:66-71). */
    static const u64 bls_r[4] = {0xFFFFFFFF00000001ULL, 0x53BDA402FFFE5BFEULL, 0x3339D80809A1D805ULL,
                                 0x73EDA753299D7D48ULL};
    /* BN254 base field p and scalar field r (bn254/cpp/bn254_fp.hpp:75-81). */
    static const u64 bn_p[4] = {0x3C208C16D87CFD47ULL, 0x97816A916871CA8DULL, 0xB85045B68181585DULL,
                                0x30644E72E131A029ULL};
    static const u64 bn_r[4] = {0x43E1F593F0000001ULL, 0x2833E84879B97091ULL, 0xB85045B68181585DULL,
                                0x30644E72E131A029ULL};
    /* Banderwagon scalar field r (banderwagon/cpp/fr.cpp:11-16). */
    static const u64 bw_r[4] = {0x74FD06B52876E7E1ULL, 0xFF8F870074190471ULL, 0x0CCE760202687600ULL,
                                0x1CFB69D4CA675F52ULL};
    memset(g_fields, 0, sizeof g_fields);
    g_fields[ORC_FIELD_BLS_FP].n = 6; memcpy(g_fields[ORC_FIELD_BLS_FP].p, bls_p, sizeof bls_p);
    g_fields[ORC_FIELD_BLS_FR].n = 4; memcpy(g_fields[ORC_FIELD_BLS_FR].p, bls_r, sizeof bls_r);
    g_fields[ORC_FIELD_BN_FP].n = 4;  memcpy(g_fields[ORC_FIELD_BN_FP].p, bn_p, sizeof bn_p);
    g_fields[ORC_FIELD_BN_FR].n = 4;  memcpy(g_fields[ORC_FIELD_BN_FR].p, bn_r, sizeof bn_r);
    g_fields[ORC_FIELD_BW_FR].n = 4;  memcpy(g_fields[ORC_FIELD_BW_FR].p, bw_r, sizeof bw_r);
    for (int i = 0; i < 5; ++i) derive(&g_fields[i]);
    g_init = 1;
}

const orc_field* orc_get_field(int id) {
    init_fields();
    return &g_fields[id];
}

/* ---- exported probes ---- */
int orc_field_nbytes(int field) { return 8 * orc_get_field(field)->n; }

void orc_field_modulus(int field, uint8_t* out) {
    const orc_field* F = orc_get_field(field);
    limbs_to_le(out, F->p, 8 * F->n);
}

void orc_field_mont_consts(int field, uint8_t* r1, uint8_t* r2, uint64_t* pinv) {
    const orc_field* F = orc_get_field(field);
    limbs_to_le(r1, F->r1, 8 * F->n);
    limbs_to_le(r2, F->r2, 8 * F->n);
    *pinv = F->pinv;
}

int orc_field_op(int field, int op, const uint8_t* a_le, const uint8_t* b_le, uint8_t* out) {
    const orc_field* F = orc_get_field(field);
    const int nb = 8 * F->n;
    u64 a[6] = {0}, b[6] = {0}, am[6], bm[6], r[6];
    le_to_limbs(a, a_le, nb);
    if (b_le) le_to_limbs(b, b_le, nb);
    /* inputs are taken mod p (to_mont reduces any value < 2^(64n)) */
    f_to_mont(F, am, a);
    f_to_mont(F, bm, b);
    switch (op) {
        case 0: f_add(F, r, am, bm); break;
        case 1: f_sub(F, r, am, bm); break;
        case 2: f_mul(F, r, am, bm); break;
        case 3:
            if (mp_is_zero(am, F->n)) return -1;
            f_inv(F, r, am);
            break;
        case 4:
            if (!f_sqrt(F, r, am)) return -1;
            break;
        default: return -1;
    }
    u64 c[6];
    f_from_mont(F, c, r);
    limbs_to_le(out, c, nb);
    return 0;
}
