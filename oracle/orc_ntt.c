/*
 * Oracle scalar-field NTT and KZG commitment.  Test infrastructure only.
 *
 * The reference has no working Fr NTT (ntt/ is 64-bit-word only; the Metal Fr
 * FFT in kzg/gpu/metal/kzg.metal:379-435 is never dispatched), so this follows
 * the loop structure of the reference's radix-2 NTT, ntt/cpp/ntt.cpp:97-172
 * (bit-reversal permutation, then log2 n DIT stages with w *= wm, inverse
 * scaled by n^-1), over the 255-bit BLS12-381 Fr and the 254-bit BN254 Fr.
 *
 * Domain: X_k = sum_j x_j w^{jk}, natural order in and out, w = g^((r-1)/n)
 * with g = 7 (BLS12-381 Fr, 2-adicity 32) or g = 5 (BN254 Fr, 2-adicity 28).
 */
#include "orc_internal.h"

#include <stdlib.h>
#include <string.h>

static const orc_field* fr_field(int fr) {
    return orc_get_field(fr == 0 ? ORC_FIELD_BLS_FR : ORC_FIELD_BN_FR);
}

static int two_adicity(int fr) { return fr == 0 ? 32 : 28; }

/* w_n in Montgomery form, or -1 if n is not a supported power of two. */
static int root_mont(int fr, size_t n, u64* w) {
    const orc_field* R = fr_field(fr);
    int lg = 0;
    while (((size_t)1 << lg) < n) ++lg;
    if (((size_t)1 << lg) != n || lg > two_adicity(fr)) return -1;
    u64 e[4], g[4] = {fr == 0 ? 7u : 5u, 0, 0, 0}, gm[4];
    u64 one[4] = {1, 0, 0, 0};
    mp_sub(e, R->p, one, 4);
    /* e = (r - 1) >> lg */
    for (int s = 0; s < lg; ++s) {
        for (int i = 0; i < 4; ++i) e[i] = (e[i] >> 1) | (i < 3 ? e[i + 1] << 63 : 0);
    }
    f_to_mont(R, gm, g);
    f_pow(R, w, gm, e, 4);
    return 0;
}

int orc_root_of_unity(int fr, size_t n, uint8_t* out) {
    const orc_field* R = fr_field(fr);
    u64 w[4], c[4];
    if (root_mont(fr, n, w)) return -1;
    f_from_mont(R, c, w);
    limbs_to_le(out, c, 32);
    return 0;
}

static int load(const orc_field* R, u64 (*a)[4], const uint8_t* in, size_t n) {
    for (size_t i = 0; i < n; ++i) {
        u64 t[4];
        le_to_limbs(t, in + 32 * i, 32);
        if (mp_cmp(t, R->p, 4) >= 0) return -1;
        f_to_mont(R, a[i], t);
    }
    return 0;
}

static void store(const orc_field* R, uint8_t* out, u64 (*a)[4], size_t n) {
    for (size_t i = 0; i < n; ++i) {
        u64 t[4];
        f_from_mont(R, t, a[i]);
        limbs_to_le(out + 32 * i, t, 32);
    }
}

int orc_ntt(int fr, uint8_t* elems, size_t n, int inverse) {
    const orc_field* R = fr_field(fr);
    if (n == 0) return -1;
    if (n == 1) {
        u64 t[4];
        le_to_limbs(t, elems, 32);
        return mp_cmp(t, R->p, 4) >= 0 ? -1 : 0;
    }
    u64 w[4];
    if (root_mont(fr, n, w)) return -1;
    if (inverse) f_inv(R, w, w);
    u64 (*a)[4] = (u64 (*)[4])malloc(sizeof(u64[4]) * n);
    if (load(R, a, elems, n)) { free(a); return -1; }
    int lg = 0;
    while (((size_t)1 << lg) < n) ++lg;
    /* bit-reversal permutation (ntt.cpp butterfly prologue) */
    for (size_t i = 0; i < n; ++i) {
        size_t j = 0;
        for (int b = 0; b < lg; ++b) j |= ((i >> b) & 1) << (lg - 1 - b);
        if (i < j) {
            u64 t[4];
            memcpy(t, a[i], 32);
            memcpy(a[i], a[j], 32);
            memcpy(a[j], t, 32);
        }
    }
    /* DIT stages: len = 2, 4, ..., n; wm = w^(n/len) */
    for (size_t len = 2; len <= n; len <<= 1) {
        u64 wm[4];
        memcpy(wm, w, 32);
        for (size_t k = len; k < n; k <<= 1) f_mul(R, wm, wm, wm);
        for (size_t i = 0; i < n; i += len) {
            u64 wj[4];
            memcpy(wj, R->r1, 32);
            for (size_t j = 0; j < len / 2; ++j) {
                u64 u[4], v[4];
                memcpy(u, a[i + j], 32);
                f_mul(R, v, a[i + j + len / 2], wj);
                f_add(R, a[i + j], u, v);
                f_sub(R, a[i + j + len / 2], u, v);
                f_mul(R, wj, wj, wm);
            }
        }
    }
    if (inverse) {
        u64 nn[4] = {n, 0, 0, 0}, nm[4], ninv[4];
        f_to_mont(R, nm, nn);
        f_inv(R, ninv, nm);
        for (size_t i = 0; i < n; ++i) f_mul(R, a[i], a[i], ninv);
    }
    store(R, elems, a, n);
    free(a);
    return 0;
}

/* O(n^2) DFT by definition, for pinning the fast transform. */
int orc_dft_naive(int fr, const uint8_t* in, uint8_t* out, size_t n, int inverse) {
    const orc_field* R = fr_field(fr);
    u64 w[4];
    if (n == 0 || root_mont(fr, n, w)) return -1;
    if (inverse) f_inv(R, w, w);
    u64 (*a)[4] = (u64 (*)[4])malloc(sizeof(u64[4]) * n);
    u64 (*b)[4] = (u64 (*)[4])malloc(sizeof(u64[4]) * n);
    if (load(R, a, in, n)) { free(a); free(b); return -1; }
    u64 wk[4];
    memcpy(wk, R->r1, 32);
    for (size_t k = 0; k < n; ++k) {
        u64 acc[4] = {0}, wjk[4];
        memcpy(wjk, R->r1, 32);
        for (size_t j = 0; j < n; ++j) {
            u64 t[4];
            f_mul(R, t, a[j], wjk);
            f_add(R, acc, acc, t);
            f_mul(R, wjk, wjk, wk);
        }
        memcpy(b[k], acc, 32);
        f_mul(R, wk, wk, w);
    }
    if (inverse) {
        u64 nn[4] = {n, 0, 0, 0}, nm[4], ninv[4];
        f_to_mont(R, nm, nn);
        f_inv(R, ninv, nm);
        for (size_t i = 0; i < n; ++i) f_mul(R, b[i], b[i], ninv);
    }
    store(R, out, b, n);
    free(a);
    free(b);
    return 0;
}

/* commit = compress(MSM(srs, iNTT(evals))): coefficients are canonical Fr,
 * i.e. exactly the 32-byte LE scalars of the MSM wire. */
int orc_kzg_commit(int curve, const uint8_t* srs, const uint8_t* evals, size_t n, uint8_t* out) {
    int fr = (curve == ORC_CURVE_BLS12_381) ? 0 : 1;
    uint8_t* coeffs = (uint8_t*)malloc(32 * n);
    memcpy(coeffs, evals, 32 * n);
    if (orc_ntt(fr, coeffs, n, 1)) { free(coeffs); return -1; }
    uint8_t wire[96];
    orc_msm_mt(curve, coeffs, srs, n, wire, 8);
    orc_g1_compress(curve, wire, out);
    free(coeffs);
    return 0;
}

int orc_kzg_commit_closed_form(int curve, const uint8_t* tau_le, const uint8_t* evals, size_t n,
                               uint8_t* out) {
    const orc_field* R = orc_get_curve(curve)->R;
    int fr = (curve == ORC_CURVE_BLS12_381) ? 0 : 1;
    uint8_t* coeffs = (uint8_t*)malloc(32 * n);
    memcpy(coeffs, evals, 32 * n);
    if (orc_ntt(fr, coeffs, n, 1)) { free(coeffs); return -1; }
    u64 tau[4], taum[4], acc[4] = {0};
    le_to_limbs(tau, tau_le, 32);
    f_to_mont(R, taum, tau);
    for (size_t i = n; i-- > 0;) {
        u64 c[4], cm[4];
        le_to_limbs(c, coeffs + 32 * i, 32);
        f_to_mont(R, cm, c);
        f_mul(R, acc, acc, taum);
        f_add(R, acc, acc, cm);
    }
    u64 v[4];
    f_from_mont(R, v, acc);
    uint8_t k[32], wire[96];
    limbs_to_le(k, v, 32);
    orc_g1_mul_gen(curve, k, wire);
    orc_g1_compress(curve, wire, out);
    free(coeffs);
    return 0;
}
