// KZG opening proofs over an SRS handle (kzg_open, include/kinet/crypto/kzg.h):
// y = f(z) and proof = [q(tau)] G, q(X) = (f(X) - y) / (X - z), for f given
// by its evaluations f_i = f(w^i) on the NTT domain (natural order).
//
// Everything stays in evaluation form until the last step:
//   inv_i = 1 / (z - w^i)                       (divstep inversions, per thread)
//   S     = sum_i f_i w^i inv_i                 (tree reduction)
//   y     = (z^n - 1) n^-1 S                    (barycentric formula)
//   q_i   = (y - f_i) inv_i = q(w^i)
// then q's coefficients by the inverse NTT and the proof by the MSM against
// the SRS (the commit path).  z must lie outside the domain.  The reference's
// kzg_commit_to_proof (c-abi/crypto.h:239; quotient by synthetic division,
// kzg/gpu/metal/kzg.metal:216-251) takes an EIP-4844 blob and its trusted
// setup; this is the same opening over this library's SRS handles.
#include <cstring>

#include "engine.h"
#include "field.h"
#include "kinet/crypto/kzg.h"
#include "kzg_internal.h"
#include "msm.h"
#include "ntt.h"

namespace kz {
bool canonical_fr(int curve, const uint8_t* le32);
int log2_exact(size_t n);

namespace {

// a^e, e given as 8 little-endian words
template <class F>
__device__ Fe<F> fr_pow(const Fe<F>& a, const uint32_t (&e)[8]) {
    Fe<F> r = fe_one<F>();
#pragma unroll 1
    for (int bit = 255; bit >= 0; --bit) {
        r = fe_mul<F>(r, r);
        if ((e[bit >> 5] >> (bit & 31)) & 1u) r = fe_mul<F>(r, a);
    }
    return r;
}

// w = g^((r - 1) / 2^logn), Montgomery form (g = 7 for BLS12-381 Fr, 5 for
// BN254 Fr: the NTT's domain)
template <class F>
__device__ Fe<F> domain_root(int logn, uint32_t g) {
    uint32_t e[8];
    fe_to_words<F>(fe_const<F>(F::P), e, 8);
    e[0] -= 1;  // r is odd
    for (int k = 0; k < logn; ++k) {
        for (int j = 0; j < 7; ++j) e[j] = (e[j] >> 1) | (e[j + 1] << 31);
        e[7] >>= 1;
    }
    Fe<F> gm = fe_zero<F>();
    gm.v[0] = g;
    return fr_pow<F>(fe_to_mont<F>(gm), e);
}

// w (Montgomery words) for the domain of 2^logn points
template <class F>
__global__ void k_open_root(int logn, uint32_t g, uint32_t* __restrict__ w_out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const Fe<F> w = domain_root<F>(logn, g);
#pragma unroll
    for (int k = 0; k < F::N; ++k) w_out[k] = w.v[k];
}

// per i: w^i, inv_i = 1 / (z - w^i), t_i = f_i w^i inv_i (Montgomery, 9 words
// each); err |= 2 when z is a domain point
template <class F>
__global__ void __launch_bounds__(128)
k_open_terms(const uint32_t* __restrict__ evals, uint32_t n, const uint32_t* __restrict__ w_words,
             const uint32_t* __restrict__ z_le, uint32_t* __restrict__ inv_out, uint32_t* __restrict__ t_out,
             int* __restrict__ err) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Fe<F> w;
#pragma unroll
    for (int k = 0; k < F::N; ++k) w.v[k] = w_words[k];
    Fe<F> wi = fe_one<F>(), base = w;
    for (uint32_t e = i; e; e >>= 1) {
        if (e & 1u) wi = fe_mul<F>(wi, base);
        base = fe_mul<F>(base, base);
    }
    const Fe<F> z = fe_to_mont<F>(fe_from_words<F>(z_le, 8));
    const Fe<F> d = fe_sub<F>(z, wi);
    if (fe_is_zero<F>(d)) atomicOr(err, 2);
    const Fe<F> inv = fe_inv<F>(d);
    const Fe<F> f = fe_to_mont<F>(fe_from_words<F>(evals + (size_t)i * 8, 8));
    const Fe<F> t = fe_mul<F>(fe_mul<F>(f, wi), inv);
#pragma unroll
    for (int k = 0; k < F::N; ++k) {
        inv_out[(size_t)i * F::N + k] = inv.v[k];
        t_out[(size_t)i * F::N + k] = t.v[k];
    }
}

// out[b] = sum of in[j] over block b's grid-stride share (Montgomery words)
template <class F>
__global__ void __launch_bounds__(256)
k_fr_sum(const uint32_t* __restrict__ in, uint32_t n, uint32_t* __restrict__ out) {
    __shared__ uint32_t sh[256 * F::N];
    Fe<F> acc = fe_zero<F>();
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        Fe<F> v;
#pragma unroll
        for (int k = 0; k < F::N; ++k) v.v[k] = in[(size_t)j * F::N + k];
        acc = fe_add<F>(acc, v);
    }
#pragma unroll
    for (int k = 0; k < F::N; ++k) sh[threadIdx.x * F::N + k] = acc.v[k];
    __syncthreads();
    for (uint32_t st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) {
            Fe<F> a, b;
#pragma unroll
            for (int k = 0; k < F::N; ++k) {
                a.v[k] = sh[threadIdx.x * F::N + k];
                b.v[k] = sh[(threadIdx.x + st) * F::N + k];
            }
            a = fe_add<F>(a, b);
#pragma unroll
            for (int k = 0; k < F::N; ++k) sh[threadIdx.x * F::N + k] = a.v[k];
        }
        __syncthreads();
    }
    if (threadIdx.x < F::N) out[(size_t)blockIdx.x * F::N + threadIdx.x] = sh[threadIdx.x];
}

// y = (z^n - 1) n^-1 S: Montgomery words at y_mont, canonical LE bytes at y_le
template <class F>
__global__ void k_open_y(const uint32_t* __restrict__ S_words, int logn, const uint32_t* __restrict__ z_le,
                         uint32_t* __restrict__ y_mont, uint32_t* __restrict__ y_le) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    Fe<F> S;
#pragma unroll
    for (int k = 0; k < F::N; ++k) S.v[k] = S_words[k];
    Fe<F> zn = fe_to_mont<F>(fe_from_words<F>(z_le, 8));
    for (int k = 0; k < logn; ++k) zn = fe_mul<F>(zn, zn);
    Fe<F> nn = fe_zero<F>();
    uint32_t nw[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    nw[logn >> 5] = 1u << (logn & 31);
    nn = fe_to_mont<F>(fe_from_words<F>(nw, 8));
    const Fe<F> y = fe_mul<F>(fe_mul<F>(fe_sub<F>(zn, fe_one<F>()), fe_inv<F>(nn)), S);
#pragma unroll
    for (int k = 0; k < F::N; ++k) y_mont[k] = y.v[k];
    fe_to_words<F>(fe_from_mont<F>(y), y_le, 8);
}

// q_i = (y - f_i) inv_i as canonical LE words (the inverse NTT's input)
template <class F>
__global__ void __launch_bounds__(256)
k_open_quotient(const uint32_t* __restrict__ evals, const uint32_t* __restrict__ inv, const uint32_t* __restrict__ y_mont,
                uint32_t n, uint32_t* __restrict__ q_out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Fe<F> y, iv;
#pragma unroll
    for (int k = 0; k < F::N; ++k) {
        y.v[k] = y_mont[k];
        iv.v[k] = inv[(size_t)i * F::N + k];
    }
    const Fe<F> f = fe_to_mont<F>(fe_from_words<F>(evals + (size_t)i * 8, 8));
    fe_to_words<F>(fe_from_mont<F>(fe_mul<F>(fe_sub<F>(y, f), iv)), q_out + (size_t)i * 8, 8);
}

template <class F>
void open_launch(const uint32_t* d_ev, uint32_t n, int logn, uint32_t g, const uint32_t* d_z, uint32_t* d_inv,
                 uint32_t* d_t, uint32_t* d_part, uint32_t* d_y, uint32_t* d_yle, uint32_t* d_w, uint32_t* d_q,
                 int* d_err, hipStream_t s) {
    hipLaunchKernelGGL(k_open_root<F>, dim3(1), dim3(64), 0, s, logn, g, d_w);
    hipLaunchKernelGGL(k_open_terms<F>, dim3(ceil_div(n, 128)), dim3(128), 0, s, d_ev, n, d_w, d_z, d_inv, d_t,
                       d_err);
    const unsigned blocks = (unsigned)std::min<size_t>(256, ceil_div(n, 256));
    hipLaunchKernelGGL(k_fr_sum<F>, dim3(blocks), dim3(256), 0, s, d_t, n, d_part);
    hipLaunchKernelGGL(k_fr_sum<F>, dim3(1), dim3(256), 0, s, d_part, blocks, d_part + 256 * F::N);
    hipLaunchKernelGGL(k_open_y<F>, dim3(1), dim3(64), 0, s, d_part + 256 * F::N, logn, d_z, d_y, d_yle);
    hipLaunchKernelGGL(k_open_quotient<F>, dim3(ceil_div(n, 256)), dim3(256), 0, s, d_ev, d_inv, d_y, n, d_q);
    KZ_HIP(hipGetLastError());
}

// ------------------------------------------------- coefficient-form opening
// f(X) = sum c_i X^i divided by (X - z): with the suffix sums
//   s_i = sum_{j >= i} c_j z^(j - i)        (s_i = c_i + z s_(i+1))
// the quotient is q_k = s_(k+1), k < n - 1, and y = f(z) = s_0 -- the
// reference's synthetic division (kzg/gpu/metal/kzg.metal:201-251: p(z) by
// Horner, q[i] = p[i+1] + z q[i+1]), here as a blocked suffix scan.  Level 0
// cuts [0, n) into segments of kSynthSeg elements; a segment's total
// T_k = sum_j c_(kL+j) z^j obeys S_k = T_k + z^L S_(k+1) for the suffix value
// S_k at its first element, the same recurrence one level up with multiplier
// z^L.  Totals go up level by level until one segment remains; suffix values
// come back down, each segment starting from the suffix value of the next
// one.  Values are 8 little-endian words, canonical, in plain form; the
// multipliers are Montgomery, so a Montgomery product with one of them is
// the plain product.
constexpr uint32_t kSynthSeg = 32;

template <class F>
__device__ __forceinline__ Fe<F> load_words8(const uint32_t* p) {
    const uint4* p4 = reinterpret_cast<const uint4*>(p);
    const uint4 a = p4[0], b = p4[1];
    const uint32_t w[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
    return fe_from_words<F>(w, 8);
}

template <class F>
__device__ __forceinline__ void store_words8(uint32_t* p, const Fe<F>& v) {
    uint32_t w[8];
    fe_to_words<F>(v, w, 8);
    uint4* p4 = reinterpret_cast<uint4*>(p);
    p4[0] = make_uint4(w[0], w[1], w[2], w[3]);
    p4[1] = make_uint4(w[4], w[5], w[6], w[7]);
}

// m_0 = z (Montgomery), m_(l+1) = m_l^kSynthSeg, l < levels
template <class F>
__global__ void k_synth_multipliers(const uint32_t* __restrict__ z_le, int levels, uint32_t* __restrict__ m_out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    Fe<F> m = fe_to_mont<F>(fe_from_words<F>(z_le, 8));
    for (int l = 0;; ++l) {
#pragma unroll
        for (int k = 0; k < F::N; ++k) m_out[l * F::N + k] = m.v[k];
        if (l == levels) break;
        for (uint32_t e = 1; e < kSynthSeg; e <<= 1) m = fe_mul<F>(m, m);
    }
}

// One thread per segment k of [0, cnt): acc starts at carry[k + 1] (0 for the
// last segment, or when carry is null) and runs down the segment,
// acc = acc m + in_j.  Totals: the segment's acc goes to out[k].  Apply: acc
// is written at every j -- out[j - shift] for j >= shift, y_out for j < shift
// (shift 1 drops s_0 into y and the rest into the quotient).
template <class F, bool Apply>
__global__ void __launch_bounds__(256)
k_synth_segments(const uint32_t* __restrict__ in, uint32_t cnt, const uint32_t* __restrict__ m_mont,
                 const uint32_t* __restrict__ carry, uint32_t* __restrict__ out, uint32_t shift,
                 uint32_t* __restrict__ y_out) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t nseg = (cnt + kSynthSeg - 1) / kSynthSeg;
    if (k >= nseg) return;
    Fe<F> m;
#pragma unroll
    for (int t = 0; t < F::N; ++t) m.v[t] = m_mont[t];
    Fe<F> acc = (carry && k + 1 < nseg) ? load_words8<F>(carry + (size_t)(k + 1) * 8) : fe_zero<F>();
    const uint32_t lo = k * kSynthSeg, hi = min(cnt, lo + kSynthSeg);
#pragma unroll 1
    for (uint32_t j = hi; j-- > lo;) {
        acc = fe_add<F>(fe_mul<F>(acc, m), load_words8<F>(in + (size_t)j * 8));
        if (Apply) {
            if (j >= shift)
                store_words8<F>(out + (size_t)(j - shift) * 8, acc);
            else
                store_words8<F>(y_out, acc);
        }
    }
    if (!Apply) store_words8<F>(out + (size_t)k * 8, acc);
}

// Element counts per level: cnt[0] = n, cnt[l + 1] = ceil(cnt[l] / kSynthSeg)
// until a level fits one segment.  Returns the top level.
int synth_levels(size_t n, uint32_t* cnt) {
    int top = 0;
    cnt[0] = (uint32_t)n;
    while (cnt[top] > kSynthSeg) {
        cnt[top + 1] = (cnt[top] + kSynthSeg - 1) / kSynthSeg;
        ++top;
    }
    return top;
}

// Scratch words of synth_launch beyond its inputs and outputs: the
// multipliers, then totals and suffix values per level above 0.
size_t synth_scratch_words(size_t n) {
    uint32_t cnt[8];
    const int top = synth_levels(n, cnt);
    size_t w = 8 * 16;  // multipliers (at most 7 x 9 words), rounded up
    for (int l = 1; l <= top; ++l) w += 2 * 8 * (size_t)cnt[l];
    return w;
}

// q (n - 1 records of 8 words) and y (8 words) from the coefficients d_c
// (n >= 1 records) and z (8 LE words); scratch of synth_scratch_words(n).
template <class F>
void synth_launch(const uint32_t* d_c, size_t n, const uint32_t* d_z, uint32_t* d_q, uint32_t* d_y,
                  uint32_t* scratch, hipStream_t s) {
    uint32_t cnt[8];
    const int top = synth_levels(n, cnt);
    uint32_t* m = scratch;
    uint32_t* tot[8] = {};
    uint32_t* suf[8] = {};
    uint32_t* p = scratch + 8 * 16;
    for (int l = 1; l <= top; ++l) {
        tot[l] = p;
        suf[l] = p + 8 * (size_t)cnt[l];
        p += 2 * 8 * (size_t)cnt[l];
    }
    hipLaunchKernelGGL(k_synth_multipliers<F>, dim3(1), dim3(64), 0, s, d_z, top, m);
    auto grid = [](uint32_t c) { return dim3(ceil_div(ceil_div(c, kSynthSeg), 256)); };
    for (int l = 0; l < top; ++l)
        hipLaunchKernelGGL((k_synth_segments<F, false>), grid(cnt[l]), dim3(256), 0, s, l ? tot[l] : d_c, cnt[l],
                           m + l * F::N, nullptr, tot[l + 1], 0u, nullptr);
    for (int l = top; l >= 0; --l)
        hipLaunchKernelGGL((k_synth_segments<F, true>), grid(cnt[l]), dim3(256), 0, s, l ? tot[l] : d_c, cnt[l],
                           m + l * F::N, l < top ? suf[l + 1] : nullptr, l ? suf[l] : d_q, l ? 0u : 1u, d_y);
    KZ_HIP(hipGetLastError());
}

}  // namespace
}  // namespace kz

using namespace kz;

extern "C" int kzg_open(const kzg_srs* srs, const uint8_t* evals, size_t n, const uint8_t z[32], uint8_t* proof_out,
                        uint8_t y_out[32]) {
    if (!srs || !evals || !z || !proof_out || !y_out) return KZ_CRYPTO_ERR_INPUT;
    const int logn = log2_exact(n);
    if (n < 2 || n > srs->n || logn < 0 || logn > (srs->curve == 2 ? 32 : 28) || n >= ((size_t)1 << 31))
        return KZ_CRYPTO_ERR_INPUT;
    if (!canonical_fr(srs->curve, z)) return KZ_CRYPTO_ERR_INPUT;
    for (size_t i = 0; i < n; ++i)
        if (!canonical_fr(srs->curve, evals + 32 * i)) return KZ_CRYPTO_ERR_INPUT;
    try {
        Engine& e = engine();
        std::lock_guard<std::mutex> g(e.mu);
        hipStream_t s = e.stream;
        constexpr size_t W = 9;  // Montgomery words of an Fr element
        uint32_t* d_ev = static_cast<uint32_t*>(e.h2d_a.ensure(32 * n));
        uint8_t* scratch = static_cast<uint8_t*>(e.ntt_a.ensure(n * (8 * W + 32 + 32) + 4096 * 2 * W * 4));
        uint32_t* d_inv = reinterpret_cast<uint32_t*>(scratch);
        uint32_t* d_t = d_inv + n * W;
        uint32_t* d_q = d_t + n * W;                  // 32-byte canonical records
        uint32_t* d_coef = d_q + n * 8;               // inverse NTT output
        uint32_t* d_part = d_coef + n * 8;            // 257 partial sums
        uint32_t* d_misc = d_part + 257 * W;          // y (W words), y LE (8), z LE (8), err
        uint32_t* d_y = d_misc, *d_yle = d_misc + 16, *d_z = d_misc + 32, *d_w = d_misc + 48;
        int* d_err = reinterpret_cast<int*>(d_misc + 64);
        uint32_t* d_tmp = static_cast<uint32_t*>(e.ntt_b.ensure(32 * n));
        uint8_t* d_out = static_cast<uint8_t*>(e.out.ensure(msm_out_bytes(srs->curve) + 16));
        KZ_HIP(hipMemcpyAsync(d_ev, evals, 32 * n, hipMemcpyHostToDevice, s));
        KZ_HIP(hipMemcpyAsync(d_z, z, 32, hipMemcpyHostToDevice, s));
        KZ_HIP(hipMemsetAsync(d_err, 0, sizeof(int), s));
        if (srs->curve == 2)
            open_launch<BlsFr>(d_ev, (uint32_t)n, logn, 7, d_z, d_inv, d_t, d_part, d_y, d_yle, d_w, d_q, d_err, s);
        else
            open_launch<BnFr>(d_ev, (uint32_t)n, logn, 5, d_z, d_inv, d_t, d_part, d_y, d_yle, d_w, d_q, d_err, s);
        ntt_device(srs->curve == 2 ? 0 : 1, logn, true, d_q, d_coef, d_tmp, d_err, s);
        srs_msm(srs, d_coef, n, 0, e.ws, s, d_out);
        uint8_t host[96 + 48];
        int err = 0;
        uint8_t yb[32];
        KZ_HIP(hipMemcpyAsync(host, d_out, msm_out_bytes(srs->curve), hipMemcpyDeviceToHost, s));
        KZ_HIP(hipMemcpyAsync(yb, d_yle, 32, hipMemcpyDeviceToHost, s));
        KZ_HIP(hipMemcpyAsync(&err, d_err, sizeof err, hipMemcpyDeviceToHost, s));
        KZ_HIP(hipStreamSynchronize(s));
        if (err) return KZ_CRYPTO_ERR_INPUT;  // z in the domain
        std::memcpy(proof_out, host + point_wire_bytes(srs->curve), compressed_bytes(srs->curve));
        std::memcpy(y_out, yb, 32);
        return KZ_CRYPTO_OK;
    } catch (const std::exception& ex) {
        set_last_error(ex.what());
        return KZ_CRYPTO_ERR_BACKEND;
    }
}

extern "C" int kzg_open_coeffs(const kzg_srs* srs, const uint8_t* coeffs, size_t n, const uint8_t z[32],
                               uint8_t* proof_out, uint8_t y_out[32]) {
    if (!srs || !coeffs || !z || !proof_out || !y_out || n == 0 || n > srs->n || n >= ((size_t)1 << 31))
        return KZ_CRYPTO_ERR_INPUT;
    if (!canonical_fr(srs->curve, z)) return KZ_CRYPTO_ERR_INPUT;
    for (size_t i = 0; i < n; ++i)
        if (!canonical_fr(srs->curve, coeffs + 32 * i)) return KZ_CRYPTO_ERR_INPUT;
    try {
        Engine& e = engine();
        std::lock_guard<std::mutex> g(e.mu);
        hipStream_t s = e.stream;
        uint32_t* d_c = static_cast<uint32_t*>(e.h2d_a.ensure(32 * n));
        // quotient (n - 1 records), then y and z (8 words each), then the scan's scratch
        uint32_t* d_q = static_cast<uint32_t*>(e.ntt_a.ensure(32 * n + 64 + 4 * synth_scratch_words(n)));
        uint32_t* d_y = d_q + 8 * (n - 1);
        uint32_t* d_z = d_y + 8;
        uint32_t* d_scr = d_z + 8;
        uint8_t* d_out = static_cast<uint8_t*>(e.out.ensure(msm_out_bytes(srs->curve) + 16));
        KZ_HIP(hipMemcpyAsync(d_c, coeffs, 32 * n, hipMemcpyHostToDevice, s));
        KZ_HIP(hipMemcpyAsync(d_z, z, 32, hipMemcpyHostToDevice, s));
        if (srs->curve == 2)
            synth_launch<BlsFr>(d_c, n, d_z, d_q, d_y, d_scr, s);
        else
            synth_launch<BnFr>(d_c, n, d_z, d_q, d_y, d_scr, s);
        if (n > 1) {
            srs_msm(srs, d_q, n - 1, 0, e.ws, s, d_out);
        } else {  // q = 0: the proof is the identity
            KZ_HIP(hipMemsetAsync(d_out, 0, msm_out_bytes(srs->curve), s));
            compress_on_device(srs->curve, d_out, d_out + point_wire_bytes(srs->curve), s);
        }
        KZ_HIP(hipMemcpyAsync(proof_out, d_out + point_wire_bytes(srs->curve), compressed_bytes(srs->curve),
                              hipMemcpyDeviceToHost, s));
        KZ_HIP(hipMemcpyAsync(y_out, d_y, 32, hipMemcpyDeviceToHost, s));
        KZ_HIP(hipStreamSynchronize(s));
        return KZ_CRYPTO_OK;
    } catch (const std::exception& ex) {
        set_last_error(ex.what());
        return KZ_CRYPTO_ERR_BACKEND;
    }
}
