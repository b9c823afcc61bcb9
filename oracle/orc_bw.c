/*
 * Banderwagon MSM oracle (test infrastructure only).
 *
 * Restates the reference's Banderwagon path behind gpukit_multi_pippenger_cpu
 * curve 3 (gpukit/cpp/cpu_reference/multi_pippenger.cpp:314-350), which reads
 * the points with Element::deserialize_uncompressed and the scalars with
 * Fr::from_bytes_le and sums with kinet::banderwagon::multi_scalar_mul
 * (banderwagon/cpp/multiexp.cpp:36-157):
 *
 *   curve     a x^2 + y^2 = 1 + d x^2 y^2 over q = BLS12-381 r, a = -5,
 *             d = 0x6389c126...188d58e7 (element.cpp:22-45)
 *   points    projective (X : Y : Z), identity (0 : 1 : 1); unified
 *             add-2008-bbjlp and dbl-2008-bbjlp (element.cpp:145-209)
 *   window    c in [4, 16] minimising (256 / c)(n + 2^c) (multiexp.cpp:40-57)
 *   digits    ceil(256 / c) windows of signed digits with a carry
 *             (multiexp.cpp:82-127); the scalars are below the scalar field
 *             order r' < 2^253, so no carry leaves the top window
 *   reduce    running sums from the top bucket (multiexp.cpp:129-136)
 *   combine   MSW -> LSW with c doublings between windows (multiexp.cpp:143-154)
 *   n = 1     Element::scalar_mul (multiexp.cpp:113-115)
 *
 * Every step is an exact group operation on the Bandersnatch curve (the
 * unified formulas have no exceptional case on the points that pass the
 * subgroup check), so the result is the exact sum sum s_i P_i of the given
 * representatives and its affine (x, y) is what serialize_uncompressed writes.
 */
#include "orc_internal.h"

#include <stdlib.h>
#include <string.h>

typedef struct {
    u64 X[4], Y[4], Z[4];
} bw_pt;

static const orc_field* Fq(void) { return orc_get_field(ORC_FIELD_BLS_FR); }

/* Curve constants in Montgomery form, from canonical big-endian values. */
static const uint8_t kD_be[32] = {0x63, 0x89, 0xc1, 0x26, 0x33, 0xc2, 0x67, 0xcb, 0xc6, 0x6e, 0x3b,
                                  0xf8, 0x6b, 0xe3, 0xb6, 0xd8, 0xcb, 0x66, 0x67, 0x71, 0x77, 0xe5,
                                  0x4f, 0x92, 0xb3, 0x69, 0xf2, 0xf5, 0x18, 0x8d, 0x58, 0xe7};
static const uint8_t kGx_be[32] = {0x29, 0xc1, 0x32, 0xcc, 0x2c, 0x0b, 0x34, 0xc5, 0x74, 0x37, 0x11,
                                   0x77, 0x7b, 0xbe, 0x42, 0xf3, 0x2b, 0x79, 0xc0, 0x22, 0xad, 0x99,
                                   0x84, 0x65, 0xe1, 0xe7, 0x18, 0x66, 0xa2, 0x52, 0xae, 0x18};
static const uint8_t kGy_be[32] = {0x2a, 0x6c, 0x66, 0x9e, 0xda, 0x12, 0x3e, 0x0f, 0x15, 0x7d, 0x8b,
                                   0x50, 0xba, 0xdc, 0xd5, 0x86, 0x35, 0x8c, 0xad, 0x81, 0xee, 0xe4,
                                   0x64, 0x60, 0x5e, 0x31, 0x67, 0xb6, 0xcc, 0x97, 0x41, 0x66};

static void mont_from_be(u64* m, const uint8_t* be) {
    u64 t[4];
    be_to_limbs(t, be, 32);
    f_to_mont(Fq(), m, t);
}

static void mont_small(u64* m, u64 v) {
    u64 t[4] = {v, 0, 0, 0};
    f_to_mont(Fq(), m, t);
}

static void const_a(u64* a) {  /* -5 */
    static u64 c[4];
    static int init;
    if (!init) {
        u64 five[4];
        mont_small(five, 5);
        f_neg(Fq(), c, five);
        init = 1;
    }
    memcpy(a, c, 32);
}

static void const_d(u64* d) {
    static u64 c[4];
    static int init;
    if (!init) {
        mont_from_be(c, kD_be);
        init = 1;
    }
    memcpy(d, c, 32);
}

static void one_m(u64* r) { memcpy(r, Fq()->r1, 32); }

static void bw_identity(bw_pt* r) {
    memset(r->X, 0, 32);
    one_m(r->Y);
    one_m(r->Z);
}

/* add-2008-bbjlp: A = Z1 Z2, B = A^2, C = X1 X2, D = Y1 Y2, E = d C D,
 * F = B - E, G = B + E, X3 = A F ((X1 + Y1)(X2 + Y2) - C - D),
 * Y3 = A G (D - a C), Z3 = F G. */
static void bw_add(bw_pt* r, const bw_pt* p, const bw_pt* q) {
    const orc_field* F = Fq();
    u64 a[4], d[4], A[4], B[4], C[4], D[4], E[4], Fv[4], G[4], H[4], I[4], t[4], u[4];
    const_a(a);
    const_d(d);
    f_mul(F, A, p->Z, q->Z);
    f_mul(F, B, A, A);
    f_mul(F, C, p->X, q->X);
    f_mul(F, D, p->Y, q->Y);
    f_mul(F, t, C, D);
    f_mul(F, E, d, t);
    f_sub(F, Fv, B, E);
    f_add(F, G, B, E);
    f_add(F, H, p->X, p->Y);
    f_add(F, I, q->X, q->Y);
    bw_pt out;
    f_mul(F, t, H, I);
    f_sub(F, t, t, C);
    f_sub(F, t, t, D);
    f_mul(F, t, t, A);
    f_mul(F, out.X, t, Fv);
    f_mul(F, u, a, C);
    f_sub(F, u, D, u);
    f_mul(F, u, u, A);
    f_mul(F, out.Y, u, G);
    f_mul(F, out.Z, Fv, G);
    *r = out;
}

/* dbl-2008-bbjlp: B = (X + Y)^2, C = X^2, D = Y^2, E = a C, F = E + D,
 * H = Z^2, J = F - 2H, X3 = (B - C - D) J, Y3 = F (E - D), Z3 = F J. */
static void bw_dbl(bw_pt* r, const bw_pt* p) {
    const orc_field* F = Fq();
    u64 a[4], B[4], C[4], D[4], E[4], Fv[4], H[4], J[4], t[4];
    const_a(a);
    f_add(F, t, p->X, p->Y);
    f_mul(F, B, t, t);
    f_mul(F, C, p->X, p->X);
    f_mul(F, D, p->Y, p->Y);
    f_mul(F, E, a, C);
    f_add(F, Fv, E, D);
    f_mul(F, H, p->Z, p->Z);
    f_add(F, t, H, H);
    f_sub(F, J, Fv, t);
    bw_pt out;
    f_sub(F, t, B, C);
    f_sub(F, t, t, D);
    f_mul(F, out.X, t, J);
    f_sub(F, t, E, D);
    f_mul(F, out.Y, Fv, t);
    f_mul(F, out.Z, Fv, J);
    *r = out;
}

static void bw_neg(bw_pt* r, const bw_pt* p) {
    *r = *p;
    f_neg(Fq(), r->X, p->X);
}

/* Exact k P for a 256-bit LE k, right to left (as Element::scalar_mul,
 * element.cpp:233-255, without the constant-time select). */
static void bw_mul(bw_pt* r, const bw_pt* p, const uint8_t* k_le) {
    bw_pt acc, base = *p;
    bw_identity(&acc);
    int top = 255;
    while (top >= 0 && !((k_le[top >> 3] >> (top & 7)) & 1)) --top;
    for (int i = 0; i <= top; ++i) {
        if ((k_le[i >> 3] >> (i & 7)) & 1) bw_add(&acc, &acc, &base);
        bw_dbl(&base, &base);
    }
    *r = acc;
}

static void bw_affine(const bw_pt* p, u64* x, u64* y) {
    const orc_field* F = Fq();
    u64 zi[4], xm[4], ym[4];
    f_inv(F, zi, p->Z);
    f_mul(F, xm, p->X, zi);
    f_mul(F, ym, p->Y, zi);
    f_from_mont(F, x, xm);
    f_from_mont(F, y, ym);
}

static void bw_write(uint8_t* out64, const bw_pt* p) {
    u64 x[4], y[4];
    bw_affine(p, x, y);
    limbs_to_be(out64, x, 32);
    limbs_to_be(out64 + 32, y, 32);
}

/* deserialize_uncompressed (element.cpp:330-355): canonical x and y,
 * d x^2 - 1 != 0, y^2 = (a x^2 - 1) / (d x^2 - 1), and 1 - a x^2 a non-zero
 * square (the subgroup test).  All-zero bytes are the identity. */
static int bw_read(bw_pt* r, const uint8_t* w) {
    const orc_field* F = Fq();
    int any = 0;
    for (int i = 0; i < 64; ++i) any |= w[i];
    if (!any) {
        bw_identity(r);
        return 1;
    }
    u64 x[4], y[4];
    be_to_limbs(x, w, 32);
    be_to_limbs(y, w + 32, 32);
    if (mp_cmp(x, F->p, 4) >= 0 || mp_cmp(y, F->p, 4) >= 0) return 0;
    u64 xm[4], ym[4], a[4], d[4], one[4], x2[4], num[4], den[4], t[4], y2[4];
    f_to_mont(F, xm, x);
    f_to_mont(F, ym, y);
    const_a(a);
    const_d(d);
    one_m(one);
    f_mul(F, x2, xm, xm);
    f_mul(F, t, a, x2);
    f_sub(F, num, t, one);
    f_mul(F, t, d, x2);
    f_sub(F, den, t, one);
    if (mp_is_zero(den, 4)) return 0;
    f_inv(F, t, den);
    f_mul(F, t, num, t);
    f_mul(F, y2, ym, ym);
    if (mp_cmp(y2, t, 4) != 0) return 0;
    /* Legendre symbol of 1 - a x^2 by Euler's criterion */
    u64 s[4], e[4], one_plain[4] = {1, 0, 0, 0}, leg[4];
    f_mul(F, t, a, x2);
    f_sub(F, s, one, t);
    if (mp_is_zero(s, 4)) return 0;
    mp_sub(e, F->p, one_plain, 4);
    for (int i = 0; i < 4; ++i) e[i] = (e[i] >> 1) | (i + 1 < 4 ? e[i + 1] << 63 : 0);
    f_pow(F, leg, s, e, 4);
    if (mp_cmp(leg, one, 4) != 0) return 0;
    memcpy(r->X, xm, 32);
    memcpy(r->Y, ym, 32);
    one_m(r->Z);
    return 1;
}

/* Fr::from_bytes_le: canonical below the scalar field order. */
static int scalar_ok(const uint8_t* s) {
    u64 t[4];
    le_to_limbs(t, s, 32);
    return mp_cmp(t, orc_get_field(ORC_FIELD_BW_FR)->p, 4) < 0;
}

static void generator(bw_pt* g) {
    mont_from_be(g->X, kGx_be);
    mont_from_be(g->Y, kGy_be);
    one_m(g->Z);
}

void orc_bw_generator(uint8_t* out64) {
    bw_pt g;
    generator(&g);
    bw_write(out64, &g);
}

int orc_bw_check(const uint8_t* p64) {
    bw_pt p;
    return bw_read(&p, p64);
}

void orc_bw_mul(const uint8_t* p64, const uint8_t* k_le, uint8_t* out64) {
    bw_pt p, r;
    bw_read(&p, p64);
    bw_mul(&r, &p, k_le);
    bw_write(out64, &r);
}

void orc_bw_add(const uint8_t* a64, const uint8_t* b64, uint8_t* out64) {
    bw_pt a, b;
    bw_read(&a, a64);
    bw_read(&b, b64);
    bw_add(&a, &a, &b);
    bw_write(out64, &a);
}

void orc_bw_dbl(const uint8_t* a64, uint8_t* out64) {
    bw_pt a;
    bw_read(&a, a64);
    bw_dbl(&a, &a);
    bw_write(out64, &a);
}

/* serialize_compressed (element.cpp:266-284): x, negated unless y > (q-1)/2. */
void orc_bw_compress(const uint8_t* p64, uint8_t* out32) {
    const orc_field* F = Fq();
    u64 x[4], y[4], h[4], one_plain[4] = {1, 0, 0, 0};
    be_to_limbs(x, p64, 32);
    be_to_limbs(y, p64 + 32, 32);
    mp_sub(h, F->p, one_plain, 4);
    for (int i = 0; i < 4; ++i) h[i] = (h[i] >> 1) | (i + 1 < 4 ? h[i + 1] << 63 : 0);
    if (mp_cmp(y, h, 4) <= 0 && !mp_is_zero(x, 4)) mp_sub(x, F->p, x, 4);
    limbs_to_be(out32, x, 32);
}

static int best_c(size_t n) {
    double best = 0;
    int c_best = 4;
    for (int c = 4; c <= 16; ++c) {
        const double cost = (256.0 / (double)c) * ((double)n + (double)(1u << c));
        if (c == 4 || cost < best) {
            best = cost;
            c_best = c;
        }
    }
    return c_best;
}

static uint32_t window_bits(const uint8_t* s, size_t off, int c) {
    uint32_t w = 0;
    const size_t b = off >> 3;
    for (int i = 0; i < 4; ++i)
        if (b + i < 32) w |= (uint32_t)s[b + i] << (8 * i);
    return (w >> (off & 7)) & ((1u << c) - 1u);
}

static int read_all(const uint8_t* scalars, const uint8_t* points, size_t n, bw_pt* pts) {
    for (size_t i = 0; i < n; ++i)
        if (!bw_read(&pts[i], points + 64 * i) || !scalar_ok(scalars + 32 * i)) return 0;
    return 1;
}

int orc_bw_msm(const uint8_t* scalars, const uint8_t* points, size_t n, uint8_t* out64) {
    bw_pt res;
    bw_identity(&res);
    if (n == 0) {
        bw_write(out64, &res);
        return 0;
    }
    bw_pt* pts = malloc(n * sizeof(bw_pt));
    if (!read_all(scalars, points, n, pts)) {
        free(pts);
        return -2;
    }
    if (n == 1) {
        bw_mul(&res, &pts[0], scalars);
        bw_write(out64, &res);
        free(pts);
        return 0;
    }
    const int c = best_c(n);
    const size_t M = (size_t)1 << (c - 1), W = (256 + (size_t)c - 1) / (size_t)c;
    bw_pt* buckets = malloc(M * sizeof(bw_pt));
    bw_pt* wsum = malloc(W * sizeof(bw_pt));
    int* carry = calloc(n, sizeof(int));
    for (size_t w = 0; w < W; ++w) {
        for (size_t k = 0; k < M; ++k) bw_identity(&buckets[k]);
        for (size_t i = 0; i < n; ++i) {
            int32_t dg = (int32_t)window_bits(scalars + 32 * i, w * (size_t)c, c) + carry[i];
            carry[i] = 0;
            if (dg > (int32_t)M) {
                dg -= (int32_t)(1u << c);
                carry[i] = 1;
            }
            if (dg > 0) {
                bw_add(&buckets[dg - 1], &buckets[dg - 1], &pts[i]);
            } else if (dg < 0) {
                bw_pt ng;
                bw_neg(&ng, &pts[i]);
                bw_add(&buckets[-dg - 1], &buckets[-dg - 1], &ng);
            }
        }
        bw_pt run, tot;
        bw_identity(&run);
        bw_identity(&tot);
        for (size_t k = M; k-- > 0;) {
            bw_add(&run, &run, &buckets[k]);
            bw_add(&tot, &tot, &run);
        }
        wsum[w] = tot;
    }
    res = wsum[W - 1];
    for (size_t w = W - 1; w-- > 0;) {
        for (int b = 0; b < c; ++b) bw_dbl(&res, &res);
        bw_add(&res, &res, &wsum[w]);
    }
    bw_write(out64, &res);
    free(carry);
    free(wsum);
    free(buckets);
    free(pts);
    return 0;
}

int orc_bw_msm_naive(const uint8_t* scalars, const uint8_t* points, size_t n, uint8_t* out64) {
    bw_pt acc, p, t;
    bw_identity(&acc);
    for (size_t i = 0; i < n; ++i) {
        if (!bw_read(&p, points + 64 * i) || !scalar_ok(scalars + 32 * i)) return -2;
        bw_mul(&t, &p, scalars + 32 * i);
        bw_add(&acc, &acc, &t);
    }
    bw_write(out64, &acc);
    return 0;
}

void orc_bw_gen_points(const uint8_t* ks, size_t n, uint8_t* out_wire) {
    bw_pt g, r;
    generator(&g);
    for (size_t i = 0; i < n; ++i) {
        bw_mul(&r, &g, ks + 32 * i);
        bw_write(out_wire + 64 * i, &r);
    }
}
