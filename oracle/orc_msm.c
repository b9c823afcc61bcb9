/*
 * Oracle MSM: the reference's signed-digit Pippenger, restated in C.
 * Test infrastructure only (see oracle.h).
 *
 *   orc_best_c         <- gpukit/cpp/cpu_reference/multi_pippenger.cpp:67-88
 *   window_digit       <- multi_pippenger.cpp:90-102
 *   process_window     <- multi_pippenger.cpp:104-144 (signed digits with carry,
 *                         bucket adds of +-P, running-sum reduction)
 *   pippenger          <- multi_pippenger.cpp:146-181 (MSW->LSW combine)
 *   orc_msm            <- gpukit_multi_pippenger_cpu, multi_pippenger.cpp:356-405
 *
 * Divergence: when c divides 256 the last window can emit a carry that the
 * reference drops; here the loop runs one more window so the carry is kept.
 */
#include "orc_internal.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>

static u64 cost_model(u64 c, size_t n) {
    const u64 windows = (256 + c - 1) / c;
    return windows * ((u64)n + ((u64)1 << c));
}

int orc_best_c(size_t n) {
    u64 best = 4, best_cost = ~(u64)0;
    for (u64 c = 4; c <= 12; ++c) {
        u64 cost = cost_model(c, n);
        if (cost < best_cost) { best_cost = cost; best = c; }
    }
    return (int)best;
}

static uint32_t window_digit(const uint8_t* s, size_t bit_offset, int c) {
    const size_t byte_idx = bit_offset >> 3;
    const size_t bit_in_byte = bit_offset & 7;
    uint32_t word = 0;
    for (int i = 0; i < 4; ++i) {
        size_t bi = byte_idx + i;
        if (bi < 32) word |= (uint32_t)s[bi] << (8 * i);
    }
    word >>= bit_in_byte;
    return word & ((1u << c) - 1u);
}

typedef struct {
    const orc_curve* C;
    const orc_jac* pts;
    const uint8_t* scalars;
    size_t n;
    int c;
    int nwin;           /* windows incl. the carry window */
    int8_t* carries;    /* per point, per window: carry INTO window w */
    orc_jac* wsum;      /* per window */
    int w_begin, w_end;
} msm_job;

/* Digits of every window for one point; carries[i*(nwin+1)+w] is the carry
 * into window w (so windows are independent once carries are known). */
static void compute_carries(msm_job* J) {
    const int c = J->c;
    const int32_t max_pos = (int32_t)1 << (c - 1);
    for (size_t i = 0; i < J->n; ++i) {
        int8_t* cr = J->carries + i * (size_t)(J->nwin + 1);
        int carry = 0;
        for (int w = 0; w < J->nwin; ++w) {
            cr[w] = (int8_t)carry;
            int32_t d = (int32_t)window_digit(J->scalars + 32 * i, (size_t)w * c, c) + carry;
            carry = 0;
            if (d > max_pos) carry = 1;
        }
        cr[J->nwin] = (int8_t)carry;
    }
}

static void process_window(msm_job* J, int w, orc_jac* buckets) {
    const orc_curve* C = J->C;
    const int c = J->c;
    const int32_t max_pos = (int32_t)1 << (c - 1);
    const size_t M = (size_t)1 << (c - 1);
    for (size_t k = 0; k < M; ++k) g_zero(&buckets[k]);
    for (size_t i = 0; i < J->n; ++i) {
        int32_t d = (int32_t)window_digit(J->scalars + 32 * i, (size_t)w * c, c) +
                    J->carries[i * (size_t)(J->nwin + 1) + w];
        if (d > max_pos) d -= (int32_t)1 << c;
        if (d == 0) continue;
        if (d > 0) {
            g_add(C, &buckets[d - 1], &buckets[d - 1], &J->pts[i]);
        } else {
            orc_jac np;
            g_neg(C, &np, &J->pts[i]);
            g_add(C, &buckets[-d - 1], &buckets[-d - 1], &np);
        }
    }
    orc_jac running, total;
    g_zero(&running);
    g_zero(&total);
    for (size_t kk = 0; kk < M; ++kk) {
        size_t k = M - 1 - kk;
        g_add(C, &running, &running, &buckets[k]);
        g_add(C, &total, &total, &running);
    }
    J->wsum[w] = total;
}

static void* window_worker(void* arg) {
    msm_job* J = (msm_job*)arg;
    orc_jac* buckets = (orc_jac*)malloc(sizeof(orc_jac) << (J->c - 1));
    for (int w = J->w_begin; w < J->w_end; ++w) process_window(J, w, buckets);
    free(buckets);
    return NULL;
}

static int msm_impl(int curve, const uint8_t* scalars, const uint8_t* points, size_t n,
                    uint8_t* out, int threads) {
    const orc_curve* C = orc_get_curve(curve);
    if (n == 0) {
        memset(out, 0, 2 * C->fp_bytes);
        return 0;
    }
    const int pb = 2 * C->fp_bytes;
    orc_jac* pts = (orc_jac*)malloc(sizeof(orc_jac) * n);
    for (size_t i = 0; i < n; ++i) g_read_wire(C, &pts[i], points + pb * i);

    msm_job J;
    memset(&J, 0, sizeof J);
    J.C = C;
    J.pts = pts;
    J.scalars = scalars;
    J.n = n;
    J.c = orc_best_c(n);
    int nwin = (256 + J.c - 1) / J.c;
    if (nwin * J.c == 256) nwin += 1;   /* room for the top carry */
    J.nwin = nwin;
    J.carries = (int8_t*)malloc(n * (size_t)(nwin + 1));
    J.wsum = (orc_jac*)malloc(sizeof(orc_jac) * nwin);
    compute_carries(&J);

    if (threads <= 1) {
        J.w_begin = 0;
        J.w_end = nwin;
        window_worker(&J);
    } else {
        if (threads > nwin) threads = nwin;
        pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * threads);
        msm_job* jobs = (msm_job*)malloc(sizeof(msm_job) * threads);
        for (int t = 0; t < threads; ++t) {
            jobs[t] = J;
            jobs[t].w_begin = (int)((long)nwin * t / threads);
            jobs[t].w_end = (int)((long)nwin * (t + 1) / threads);
            pthread_create(&th[t], NULL, window_worker, &jobs[t]);
        }
        for (int t = 0; t < threads; ++t) pthread_join(th[t], NULL);
        free(th);
        free(jobs);
    }

    orc_jac result = J.wsum[nwin - 1];
    for (int w = nwin - 2; w >= 0; --w) {
        for (int b = 0; b < J.c; ++b) g_dbl(C, &result, &result);
        g_add(C, &result, &result, &J.wsum[w]);
    }
    g_write_wire(C, out, &result);
    free(pts);
    free(J.carries);
    free(J.wsum);
    return 0;
}

int orc_msm(int curve, const uint8_t* scalars, const uint8_t* points, size_t n, uint8_t* out) {
    return msm_impl(curve, scalars, points, n, out, 1);
}

int orc_msm_mt(int curve, const uint8_t* scalars, const uint8_t* points, size_t n, uint8_t* out,
               int threads) {
    return msm_impl(curve, scalars, points, n, out, threads);
}

int orc_msm_naive(int curve, const uint8_t* scalars, const uint8_t* points, size_t n, uint8_t* out) {
    const orc_curve* C = orc_get_curve(curve);
    const int pb = 2 * C->fp_bytes;
    orc_jac acc, p, t;
    g_zero(&acc);
    for (size_t i = 0; i < n; ++i) {
        u64 k[4];
        le_to_limbs(k, scalars + 32 * i, 32);
        g_read_wire(C, &p, points + pb * i);
        g_mul(C, &t, &p, k);
        g_add(C, &acc, &acc, &t);
    }
    g_write_wire(C, out, &acc);
    return 0;
}

/* ---- seeded inputs ---- */
uint64_t orc_splitmix64(uint64_t* s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

/* Candidate = 4 consecutive SplitMix64 outputs as LE limbs, top limb masked to
 * the bit length of r; accepted when < r. */
void orc_gen_scalars(int curve, uint64_t seed, size_t n, uint8_t* out) {
    const orc_field* R = orc_get_curve(curve)->R;
    const int top_bits = R->bits - 192;
    const u64 mask = (top_bits >= 64) ? ~(u64)0 : (((u64)1 << top_bits) - 1);
    uint64_t s = seed;
    size_t i = 0;
    while (i < n) {
        u64 k[4];
        for (int j = 0; j < 4; ++j) k[j] = orc_splitmix64(&s);
        k[3] &= mask;
        if (mp_cmp(k, R->p, 4) >= 0) continue;
        limbs_to_le(out + 32 * i, k, 32);
        ++i;
    }
}

void orc_powers(int curve, const uint8_t* tau_le, size_t n, uint8_t* out) {
    const orc_field* R = orc_get_curve(curve)->R;
    u64 t[4], tm[4], acc[4], c[4];
    le_to_limbs(t, tau_le, 32);
    f_to_mont(R, tm, t);
    memcpy(acc, R->r1, 32);
    for (size_t i = 0; i < n; ++i) {
        f_from_mont(R, c, acc);
        limbs_to_le(out + 32 * i, c, 32);
        f_mul(R, acc, acc, tm);
    }
}

void orc_gen_points(int curve, const uint8_t* ks, size_t n, uint8_t* out) {
    const int pb = orc_point_bytes(curve);
    for (size_t i = 0; i < n; ++i) orc_g1_mul_gen(curve, ks + 32 * i, out + pb * i);
}

/* (k0 + i*d) * G for i < n: one Jacobian addition per point and a Montgomery
 * batch inversion (prefix products, one inverse, backward sweep) for the
 * affine normalisation, so large test sets are cheap.  A point that hits
 * infinity is written as all-zero bytes. */
void orc_gen_points_progression(int curve, const uint8_t* k0_le, const uint8_t* d_le, size_t n, uint8_t* out) {
    const orc_curve* C = orc_get_curve(curve);
    const orc_field* F = C->F;
    const int fb = C->fp_bytes;
    const int nl = F->n;
    orc_jac g, cur, step;
    u64 k0[4], d[4];
    le_to_limbs(k0, k0_le, 32);
    le_to_limbs(d, d_le, 32);
    g_generator(C, &g, curve);
    g_mul(C, &cur, &g, k0);
    g_mul(C, &step, &g, d);
    orc_jac* pts = (orc_jac*)malloc(sizeof(orc_jac) * n);
    u64 (*pref)[6] = (u64 (*)[6])malloc(sizeof(u64[6]) * (n + 1));
    memcpy(pref[0], F->r1, 8 * nl);
    for (size_t i = 0; i < n; ++i) {
        pts[i] = cur;
        const u64* z = (cur.inf || mp_is_zero(cur.Z, nl)) ? F->r1 : cur.Z;
        f_mul(F, pref[i + 1], pref[i], z);
        g_add(C, &cur, &cur, &step);
    }
    u64 inv[6];
    f_inv(F, inv, pref[n]);
    for (size_t i = n; i-- > 0;) {
        u64 zi[6], t[6];
        const int skip = pts[i].inf || mp_is_zero(pts[i].Z, nl);
        const u64* z = skip ? F->r1 : pts[i].Z;
        f_mul(F, zi, inv, pref[i]);   /* 1/z_i */
        f_mul(F, inv, inv, z);        /* drop z_i from the running inverse */
        uint8_t* w = out + (size_t)2 * fb * i;
        if (skip) { memset(w, 0, 2 * fb); continue; }
        u64 zi2[6], x[6], y[6], xc[6], yc[6];
        f_mul(F, zi2, zi, zi);
        f_mul(F, x, pts[i].X, zi2);
        f_mul(F, t, zi2, zi);
        f_mul(F, y, pts[i].Y, t);
        f_from_mont(F, xc, x);
        f_from_mont(F, yc, y);
        limbs_to_be(w, xc, fb);
        limbs_to_be(w + fb, yc, fb);
    }
    free(pts);
    free(pref);
}
