/*
 * Closed forms over the scalar field, O(n) each.  Test infrastructure only:
 * they check the device MSM, NTT and KZG paths at full size without running a
 * CPU Pippenger or NTT.
 *
 *   orc_fr_dot         sum a_i b_i mod r  (MSM over points b_i G is (dot) G)
 *   orc_fr_eval        sum c_i x^i by Horner (one output of an NTT: X_k = f(w^k))
 *   orc_fr_lagrange    f(tau) from the evaluations f(w^j), j < n, by the
 *                      barycentric formula
 *                         f(tau) = (tau^n - 1) / n * sum_j f(w^j) w^j / (tau - w^j)
 *                      with one field inversion (Montgomery's batch trick);
 *                      it uses neither an NTT nor the coefficients.
 *   orc_fr_divide_linear  f = (X - z) q + y by synthetic division (the
 *                      quotient of an opening proof from coefficients)
 * Elements are 32-byte little-endian.  Inputs of dot may be any 256-bit
 * values (reduced mod r first); eval / lagrange inputs must be canonical.
 */
#include "orc_internal.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

static const orc_field* fr_of(int fr) {
    return orc_get_field(fr == 0 ? ORC_FIELD_BLS_FR : fr == 1 ? ORC_FIELD_BN_FR : ORC_FIELD_BW_FR);
}

/* 256-bit LE value mod r, into Montgomery form */
static void load_reduced(const orc_field* R, u64* m, const uint8_t* le) {
    u64 t[4];
    le_to_limbs(t, le, 32);
    while (mp_cmp(t, R->p, 4) >= 0) mp_sub(t, t, R->p, 4);
    f_to_mont(R, m, t);
}

static int load_canonical(const orc_field* R, u64* m, const uint8_t* le) {
    u64 t[4];
    le_to_limbs(t, le, 32);
    if (mp_cmp(t, R->p, 4) >= 0) return -1;
    f_to_mont(R, m, t);
    return 0;
}

static void store_plain(const orc_field* R, uint8_t* out, const u64* m) {
    u64 t[4];
    f_from_mont(R, t, m);
    limbs_to_le(out, t, 32);
}

int orc_fr_dot(int fr, const uint8_t* a, const uint8_t* b, size_t n, uint8_t* out) {
    const orc_field* R = fr_of(fr);
    u64 acc[4] = {0, 0, 0, 0};
    for (size_t i = 0; i < n; ++i) {
        u64 x[4], y[4], p[4];
        load_reduced(R, x, a + 32 * i);
        load_reduced(R, y, b + 32 * i);
        f_mul(R, p, x, y);
        f_add(R, acc, acc, p);
    }
    store_plain(R, out, acc);
    return 0;
}

int orc_fr_eval(int fr, const uint8_t* coeffs, size_t n, const uint8_t* x_le, uint8_t* out) {
    const orc_field* R = fr_of(fr);
    u64 x[4], acc[4] = {0, 0, 0, 0};
    if (load_canonical(R, x, x_le)) return -1;
    for (size_t i = n; i-- > 0;) {
        u64 c[4];
        if (load_canonical(R, c, coeffs + 32 * i)) return -1;
        f_mul(R, acc, acc, x);
        f_add(R, acc, acc, c);
    }
    store_plain(R, out, acc);
    return 0;
}

/* Synthetic division of f = sum c_i X^i by (X - z), following the reference's
 * poly_quotient (kzg/gpu/metal/kzg.metal:216-251): q[n-2] = c[n-1],
 * q[i] = c[i+1] + z q[i+1] from the top, and y = f(z) = c[0] + z q[0] (its
 * Horner p(z), kzg.metal:201-213).  q_out receives n - 1 elements. */
int orc_fr_divide_linear(int fr, const uint8_t* coeffs, size_t n, const uint8_t* z_le, uint8_t* q_out,
                         uint8_t* y_out) {
    const orc_field* R = fr_of(fr);
    u64 z[4], acc[4] = {0, 0, 0, 0};
    if (n == 0 || load_canonical(R, z, z_le)) return -1;
    for (size_t i = n; i-- > 0;) {
        u64 c[4];
        if (load_canonical(R, c, coeffs + 32 * i)) return -1;
        f_mul(R, acc, acc, z);
        f_add(R, acc, acc, c);
        store_plain(R, i ? q_out + 32 * (i - 1) : y_out, acc);
    }
    return 0;
}

/* Partial sum  sum_{j in [j0, j1)} e_j w^j / (tau - w^j)  with one batch
 * inversion over the range (Montgomery form in *acc).  -1 if tau is a domain
 * point or an evaluation is not canonical. */
typedef struct {
    const orc_field* R;
    const uint8_t* evals;
    u64 w[4], tau[4];
    size_t j0, j1;
    u64 acc[4];
    int rc;
} lag_part;

static void* lagrange_part(void* arg) {
    lag_part* P = (lag_part*)arg;
    const orc_field* R = P->R;
    const size_t m = P->j1 - P->j0;
    P->rc = 0;
    memset(P->acc, 0, sizeof P->acc);
    if (m == 0) return NULL;
    u64(*d)[4] = (u64(*)[4])malloc(sizeof(u64[4]) * m);   /* tau - w^j */
    u64(*pre)[4] = (u64(*)[4])malloc(sizeof(u64[4]) * m); /* prefix products */
    if (!d || !pre) {
        free(d);
        free(pre);
        P->rc = -1;
        return NULL;
    }
    u64 wj[4], e0[4] = {(u64)P->j0, 0, 0, 0}, run[4];
    f_pow(R, wj, P->w, e0, 1); /* w^j0 */
    memcpy(run, R->r1, sizeof run);
    for (size_t t = 0; t < m; ++t) {
        f_sub(R, d[t], P->tau, wj);
        if (mp_is_zero(d[t], 4)) {
            P->rc = -1;
            goto out;
        }
        memcpy(pre[t], run, sizeof run);
        f_mul(R, run, run, d[t]);
        f_mul(R, wj, wj, P->w);
    }
    {
        u64 inv[4], winv[4], wk[4];
        f_inv(R, inv, run);
        f_inv(R, winv, P->w);
        f_mul(R, wk, wj, winv); /* w^(j1 - 1) */
        for (size_t t = m; t-- > 0;) {
            u64 dinv[4], e[4], v[4];
            f_mul(R, dinv, inv, pre[t]);
            f_mul(R, inv, inv, d[t]);
            if (load_canonical(R, e, P->evals + 32 * (P->j0 + t))) {
                P->rc = -1;
                goto out;
            }
            f_mul(R, v, e, wk);
            f_mul(R, v, v, dinv);
            f_add(R, P->acc, P->acc, v);
            f_mul(R, wk, wk, winv);
        }
    }
out:
    free(d);
    free(pre);
    return NULL;
}

int orc_fr_lagrange(int fr, const uint8_t* evals, size_t n, const uint8_t* tau_le, uint8_t* out) {
    const orc_field* R = fr_of(fr);
    uint8_t wb[32];
    if (n == 0 || orc_root_of_unity(fr, n, wb)) return -1;
    u64 w[4], tau[4];
    if (load_canonical(R, w, wb) || load_canonical(R, tau, tau_le)) return -1;
    enum { kMaxThreads = 16 };
    lag_part parts[kMaxThreads];
    pthread_t th[kMaxThreads];
    long ncpu = sysconf(_SC_NPROCESSORS_ONLN);
    int T = (int)(ncpu < 1 ? 1 : ncpu > kMaxThreads ? kMaxThreads : ncpu);
    if ((size_t)T > n / 4096 + 1) T = (int)(n / 4096 + 1);
    for (int k = 0; k < T; ++k) {
        parts[k].R = R;
        parts[k].evals = evals;
        memcpy(parts[k].w, w, sizeof w);
        memcpy(parts[k].tau, tau, sizeof tau);
        parts[k].j0 = n * (size_t)k / (size_t)T;
        parts[k].j1 = n * (size_t)(k + 1) / (size_t)T;
    }
    int started = 0;
    for (int k = 1; k < T; ++k)
        if (pthread_create(&th[k], NULL, lagrange_part, &parts[k]) == 0) started = k;
        else break;
    lagrange_part(&parts[0]);
    for (int k = 1; k <= started; ++k) pthread_join(th[k], NULL);
    for (int k = started + 1; k < T; ++k) lagrange_part(&parts[k]); /* threads that did not start */
    u64 acc[4] = {0, 0, 0, 0};
    for (int k = 0; k < T; ++k) {
        if (parts[k].rc) return -1;
        f_add(R, acc, acc, parts[k].acc);
    }
    /* (tau^n - 1) / n */
    u64 tn[4], ninv[4], nn[4] = {(u64)n, 0, 0, 0}, nm[4], f[4];
    memcpy(tn, tau, sizeof tn);
    for (size_t k = 1; k < n; k <<= 1) f_mul(R, tn, tn, tn);
    f_sub(R, tn, tn, R->r1);
    f_to_mont(R, nm, nn);
    f_inv(R, ninv, nm);
    f_mul(R, f, tn, ninv);
    f_mul(R, acc, acc, f);
    store_plain(R, out, acc);
    return 0;
}
