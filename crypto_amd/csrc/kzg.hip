// KZG commitments C = MSM(SRS, iNTT(evals)) on the device, and the SRS /
// context handles of include/kinet/crypto/kzg.h.
#include <cstdlib>
#include <cstring>
#include <memory>
#include <vector>

#include "engine.h"
#include "field.h"
#include "kinet/crypto/hip_runtime.h"
#include "kinet/crypto/kzg.h"
#include "kzg_internal.h"
#include "msm.h"
#include "ntt.h"
#include "glv.h"
#include "scan.h"

namespace kz {
void fixed_base_mul(int curve, const uint32_t* d_scalars, size_t n, uint8_t* d_out_wire, hipStream_t s);
}


namespace kz {

// out[i] = tau^i mod r (canonical, 8 LE words), i < n.
template <class F>
__global__ void __launch_bounds__(256) k_powers(const uint32_t* __restrict__ tau_words, uint32_t n,
                                                uint32_t* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Fe<F> base = fe_to_mont<F>(fe_from_words<F>(tau_words, 8)), acc = fe_one<F>();
    for (uint32_t e = i; e; e >>= 1) {
        if (e & 1u) acc = fe_mul<F>(acc, base);
        base = fe_mul<F>(base, base);
    }
    fe_to_words<F>(fe_from_mont<F>(acc), out + (size_t)i * 8, 8);
}

bool canonical_fr(int curve, const uint8_t* le32) {
    // r of BLS12-381 / BN254 as big-endian bytes
    static const uint8_t r_bls[32] = {0x73, 0xED, 0xA7, 0x53, 0x29, 0x9D, 0x7D, 0x48, 0x33, 0x39, 0xD8,
                                      0x08, 0x09, 0xA1, 0xD8, 0x05, 0x53, 0xBD, 0xA4, 0x02, 0xFF, 0xFE,
                                      0x5B, 0xFE, 0xFF, 0xFF, 0xFF, 0xFF, 0x00, 0x00, 0x00, 0x01};
    static const uint8_t r_bn[32] = {0x30, 0x64, 0x4E, 0x72, 0xE1, 0x31, 0xA0, 0x29, 0xB8, 0x50, 0x45,
                                     0xB6, 0x81, 0x81, 0x58, 0x5D, 0x28, 0x33, 0xE8, 0x48, 0x79, 0xB9,
                                     0x70, 0x91, 0x43, 0xE1, 0xF5, 0x93, 0xF0, 0x00, 0x00, 0x01};
    const uint8_t* r = curve == 2 ? r_bls : r_bn;
    for (int k = 0; k < 32; ++k) {
        const uint8_t b = le32[31 - k];
        if (b != r[k]) return b < r[k];
    }
    return false;
}

int log2_exact(size_t n) {
    int l = 0;
    while (((size_t)1 << l) < n) ++l;
    return ((size_t)1 << l) == n ? l : -1;
}

static void srs_upload(kzg_srs* srs, const uint8_t* d_wire_src, bool src_on_host, hipStream_t s) {
    const size_t pb = point_wire_bytes(srs->curve);
    srs->wire.ensure(pb * srs->n);
    srs->pts.ensure((size_t)affine_words(srs->curve) * 4 * srs->n);
    srs->inf.ensure(srs->n);
    if (src_on_host)
        KZ_HIP(hipMemcpyAsync(srs->wire.ptr, d_wire_src, pb * srs->n, hipMemcpyHostToDevice, s));
    points_from_wire(srs->curve, srs->wire.as<uint8_t>(), srs->pts.as<uint32_t>(), srs->inf.as<uint8_t>(), srs->n,
                     s);
}

int srs_table_near(const kzg_srs* srs, int c, int max_gap) {
    int best = -1;
    for (size_t k = 0; k < srs->table_c.size(); ++k) {
        const int gap = std::abs(srs->table_c[k] - c);
        if (gap > max_gap) continue;
        if (best < 0 || gap < std::abs(srs->table_c[best] - c) ||
            (gap == std::abs(srs->table_c[best] - c) && srs->table_c[k] > srs->table_c[best]))
            best = (int)k;
    }
    return best;
}

void srs_msm(const kzg_srs* srs, const uint32_t* d_scalars, size_t n, int window_bits, MsmWorkspace& ws,
             hipStream_t s, uint8_t* d_out) {
    MsmParams prm;
    const int k = srs_table_near(srs, window_bits ? window_bits : msm_table_window_bits(n), window_bits ? 0 : 64);
    if (k >= 0) {
        prm.c = srs->table_c[k];
        prm.table_stride = (uint32_t)srs->n;
        const uint32_t* tab = srs->tables.as<uint32_t>() + (size_t)srs->table_base[k] * affine_words(srs->curve);
        msm_device(srs->curve, tab, srs->inf.as<uint8_t>(), d_scalars, n, prm, ws, s, d_out);
        return;
    }
    prm.g1 = srs->g1;
    prm.c = window_bits;
    msm_device(srs->curve, srs->pts.as<uint32_t>(), srs->inf.as<uint8_t>(), d_scalars, n, prm, ws, s, d_out);
}

// iNTT of the evaluations into ctx->coeffs, then the MSM against the SRS.
static void commit_device(kzg_ctx* ctx, const kzg_srs* srs, const uint32_t* d_evals, size_t n, int window_bits,
                          uint8_t* d_out, int* d_err) {
    const int logn = log2_exact(n);
    uint32_t* coeffs = static_cast<uint32_t*>(ctx->coeffs.ensure(32 * n));
    uint32_t* tmp = static_cast<uint32_t*>(ctx->tmp.ensure(32 * n));
    const int field = srs->curve == 2 ? 0 : 1;
    ntt_device(field, logn, true, d_evals, coeffs, tmp, d_err, ctx->stream);
    srs_msm(srs, coeffs, n, window_bits, ctx->ws, ctx->stream, d_out);
}

static kzg_ctx* default_ctx() {
    Engine& e = engine();
    static kzg_ctx* ctx = [&e] {
        kzg_ctx* c = new kzg_ctx();
        c->stream = e.stream;
        return c;
    }();
    return ctx;
}

}  // namespace kz

using namespace kz;

extern "C" {

int kzg_srs_create(uint32_t curve, const uint8_t* points, size_t n, kzg_srs** out) {
    if (!out || (curve != 1 && curve != 2) || n == 0 || !points) return KZ_CRYPTO_ERR_INPUT;
    *out = nullptr;
    try {
        Engine& e = engine();
        std::lock_guard<std::mutex> g(e.mu);
        auto srs = std::make_unique<kzg_srs>();
        srs->curve = (int)curve;
        srs->n = n;
        srs_upload(srs.get(), points, true, e.stream);
        // membership in G1 licenses the GLV split of the MSMs against this SRS
        int* d_bad = static_cast<int*>(e.out.ensure(sizeof(int)));
        device_fill_u32(reinterpret_cast<uint32_t*>(d_bad), 0, 1, e.stream);
        check_g1(srs->curve, srs->pts.as<uint32_t>(), srs->inf.as<uint8_t>(), n, d_bad, e.stream);
        int bad = 1;
        KZ_HIP(hipMemcpyAsync(&bad, d_bad, sizeof(int), hipMemcpyDeviceToHost, e.stream));
        KZ_HIP(hipStreamSynchronize(e.stream));
        srs->g1 = bad == 0;
        *out = srs.release();
        return KZ_CRYPTO_OK;
    } catch (const std::exception& ex) {
        set_last_error(ex.what());
        return KZ_CRYPTO_ERR_BACKEND;
    }
}

int kzg_srs_generate(uint32_t curve, const uint8_t tau[32], size_t n, kzg_srs** out) {
    if (!out || !tau || (curve != 1 && curve != 2) || n == 0 || n > ((size_t)1 << 28)) return KZ_CRYPTO_ERR_INPUT;
    if (!canonical_fr((int)curve, tau)) return KZ_CRYPTO_ERR_INPUT;
    *out = nullptr;
    try {
        Engine& e = engine();
        hipStream_t s = e.stream;
        auto srs = std::make_unique<kzg_srs>();
        srs->curve = (int)curve;
        srs->n = n;
        DevBuf d_tau, d_pow;
        uint32_t tw[8];
        std::memcpy(tw, tau, 32);
        d_tau.ensure(32);
        d_pow.ensure(32 * n);
        KZ_HIP(hipMemcpyAsync(d_tau.ptr, tw, 32, hipMemcpyHostToDevice, s));
        if (curve == 2)
            hipLaunchKernelGGL(k_powers<BlsFr>, dim3(ceil_div(n, 256)), dim3(256), 0, s, d_tau.as<uint32_t>(),
                               (uint32_t)n, d_pow.as<uint32_t>());
        else
            hipLaunchKernelGGL(k_powers<BnFr>, dim3(ceil_div(n, 256)), dim3(256), 0, s, d_tau.as<uint32_t>(),
                               (uint32_t)n, d_pow.as<uint32_t>());
        KZ_HIP(hipGetLastError());
        srs->wire.ensure(point_wire_bytes((int)curve) * n);
        fixed_base_mul((int)curve, d_pow.as<uint32_t>(), n, srs->wire.as<uint8_t>(), s);
        {
            std::lock_guard<std::mutex> g(e.mu);
            srs_upload(srs.get(), nullptr, false, s);
        }
        srs->g1 = true;  // tau^i G lies in G1 by construction
        KZ_HIP(hipStreamSynchronize(s));
        *out = srs.release();
        return KZ_CRYPTO_OK;
    } catch (const std::exception& ex) {
        set_last_error(ex.what());
        return KZ_CRYPTO_ERR_BACKEND;
    }
}

void kzg_srs_destroy(kzg_srs* srs) { delete srs; }

void kzg_srs_free(kzg_srs* srs) { delete srs; }

size_t kzg_srs_size(const kzg_srs* srs) { return srs ? srs->n : 0; }

int kzg_srs_precompute(kzg_srs* srs, int window_bits) {
    if (!srs || window_bits < 0 || (window_bits > 0 && window_bits < 5) || window_bits > 18)
        return KZ_CRYPTO_ERR_INPUT;
    try {
        Engine& e = engine();
        std::lock_guard<std::mutex> g(e.mu);
        const int c = window_bits ? window_bits : msm_table_window_bits(srs->n);
        for (int have : srs->table_c)
            if (have == c) return KZ_CRYPTO_OK;
        std::vector<int> cs = srs->table_c;
        cs.push_back(c);
        uint64_t records = 0;
        std::vector<uint32_t> base;
        for (int k : cs) {
            base.push_back((uint32_t)records);
            records += (uint64_t)msm_table_rows(k) * srs->n;
        }
        if (records >= 0x80000000ull) return KZ_CRYPTO_ERR_INPUT;  // entry indices are 31-bit
        // rebuild every table into one buffer (the old one is released first)
        srs->table_c.clear();
        srs->table_base.clear();
        const size_t rec = (size_t)affine_words(srs->curve);
        srs->tables.ensure(records * rec * sizeof(uint32_t));
        for (size_t k = 0; k < cs.size(); ++k)
            points_table_build(srs->curve, srs->pts.as<uint32_t>(), srs->inf.as<uint8_t>(), srs->n, cs[k],
                               srs->tables.as<uint32_t>() + (size_t)base[k] * rec, e.stream);
        KZ_HIP(hipStreamSynchronize(e.stream));
        srs->table_c = cs;
        srs->table_base = base;
        return KZ_CRYPTO_OK;
    } catch (const std::exception& ex) {
        set_last_error(ex.what());
        return KZ_CRYPTO_ERR_BACKEND;
    }
}

size_t kzg_srs_table_bytes(const kzg_srs* srs) {
    if (!srs) return 0;
    size_t records = 0;
    for (int c : srs->table_c) records += (size_t)msm_table_rows(c) * srs->n;
    return records * affine_words(srs->curve) * sizeof(uint32_t);
}

uint32_t kzg_srs_curve(const kzg_srs* srs) { return srs ? (uint32_t)srs->curve : 0; }

int kzg_srs_export(const kzg_srs* srs, size_t first, size_t count, uint8_t* points_out) {
    if (!srs || !points_out || first + count > srs->n) return KZ_CRYPTO_ERR_INPUT;
    try {
        const size_t pb = point_wire_bytes(srs->curve);
        KZ_HIP(hipMemcpy(points_out, srs->wire.as<uint8_t>() + pb * first, pb * count, hipMemcpyDeviceToHost));
        return KZ_CRYPTO_OK;
    } catch (const std::exception& ex) {
        set_last_error(ex.what());
        return KZ_CRYPTO_ERR_BACKEND;
    }
}

static int commit_host(const kzg_srs* srs, const uint8_t* evals, size_t n, uint8_t* commit_out) {
    if (!srs || !evals || !commit_out) return KZ_CRYPTO_ERR_INPUT;
    if (n == 0 || n > srs->n || log2_exact(n) < 0 || log2_exact(n) > (srs->curve == 2 ? 32 : 28))
        return KZ_CRYPTO_ERR_INPUT;
    try {
        Engine& e = engine();
        std::lock_guard<std::mutex> g(e.mu);
        kzg_ctx* ctx = default_ctx();
        hipStream_t s = ctx->stream;
        uint32_t* d_ev = static_cast<uint32_t*>(e.h2d_a.ensure(32 * n));
        uint8_t* d_out = static_cast<uint8_t*>(e.out.ensure(msm_out_bytes(srs->curve) + 16));
        int* d_err = reinterpret_cast<int*>(d_out + msm_out_bytes(srs->curve));
        KZ_HIP(hipMemsetAsync(d_err, 0, sizeof(int), s));
        KZ_HIP(hipMemcpyAsync(d_ev, evals, 32 * n, hipMemcpyHostToDevice, s));
        commit_device(ctx, srs, d_ev, n, 0, d_out, d_err);
        uint8_t host[96 + 48 + 16];
        KZ_HIP(hipMemcpyAsync(host, d_out, msm_out_bytes(srs->curve) + sizeof(int), hipMemcpyDeviceToHost, s));
        KZ_HIP(hipStreamSynchronize(s));
        int err;
        std::memcpy(&err, host + msm_out_bytes(srs->curve), sizeof(int));
        if (err) return KZ_CRYPTO_ERR_INPUT;
        std::memcpy(commit_out, host + point_wire_bytes(srs->curve), compressed_bytes(srs->curve));
        return KZ_CRYPTO_OK;
    } catch (const std::exception& ex) {
        set_last_error(ex.what());
        return KZ_CRYPTO_ERR_BACKEND;
    }
}

int kzg_commit(const kzg_srs* srs, const uint8_t* evals, size_t n, uint8_t* commit_out) {
    return commit_host(srs, evals, n, commit_out);
}

int kzg_msm(const kzg_srs* srs, const uint8_t* scalars, size_t n, uint8_t* result) {
    if (!srs || !result || (n > 0 && !scalars) || n > srs->n) return KZ_CRYPTO_ERR_INPUT;
    try {
        Engine& e = engine();
        std::lock_guard<std::mutex> g(e.mu);
        const size_t pb = point_wire_bytes(srs->curve);
        if (n == 0) {
            std::memset(result, 0, pb);
            return KZ_CRYPTO_OK;
        }
        hipStream_t s = e.stream;
        uint32_t* d_sc = static_cast<uint32_t*>(e.h2d_a.ensure(32 * n));
        uint8_t* d_out = static_cast<uint8_t*>(e.out.ensure(msm_out_bytes(srs->curve) + 16));
        KZ_HIP(hipMemcpyAsync(d_sc, scalars, 32 * n, hipMemcpyHostToDevice, s));
        srs_msm(srs, d_sc, n, 0, e.ws, s, d_out);
        KZ_HIP(hipMemcpyAsync(result, d_out, pb, hipMemcpyDeviceToHost, s));
        KZ_HIP(hipStreamSynchronize(s));
        return KZ_CRYPTO_OK;
    } catch (const std::exception& ex) {
        set_last_error(ex.what());
        return KZ_CRYPTO_ERR_BACKEND;
    }
}

int kzg_commit_coeffs(const kzg_srs* srs, const uint8_t* coeffs, size_t n, uint8_t* commit_out) {
    if (!srs || !coeffs || !commit_out || n == 0 || n > srs->n) return KZ_CRYPTO_ERR_INPUT;
    for (size_t i = 0; i < n; ++i)
        if (!canonical_fr(srs->curve, coeffs + 32 * i)) return KZ_CRYPTO_ERR_INPUT;
    try {
        Engine& e = engine();
        std::lock_guard<std::mutex> g(e.mu);
        hipStream_t s = e.stream;
        uint32_t* d_sc = static_cast<uint32_t*>(e.h2d_a.ensure(32 * n));
        uint8_t* d_out = static_cast<uint8_t*>(e.out.ensure(msm_out_bytes(srs->curve) + 16));
        KZ_HIP(hipMemcpyAsync(d_sc, coeffs, 32 * n, hipMemcpyHostToDevice, s));
        srs_msm(srs, d_sc, n, 0, e.ws, s, d_out);
        KZ_HIP(hipMemcpyAsync(commit_out, d_out + point_wire_bytes(srs->curve), compressed_bytes(srs->curve),
                              hipMemcpyDeviceToHost, s));
        KZ_HIP(hipStreamSynchronize(s));
        return KZ_CRYPTO_OK;
    } catch (const std::exception& ex) {
        set_last_error(ex.what());
        return KZ_CRYPTO_ERR_BACKEND;
    }
}

int kzg_ctx_create(void* stream, kzg_ctx** out) {
    if (!out) return KZ_CRYPTO_ERR_INPUT;
    try {
        (void)engine();
        kzg_ctx* c = new kzg_ctx();
        c->stream = static_cast<hipStream_t>(stream);
        *out = c;
        return KZ_CRYPTO_OK;
    } catch (const std::exception& ex) {
        set_last_error(ex.what());
        return KZ_CRYPTO_ERR_BACKEND;
    }
}

void kzg_ctx_destroy(kzg_ctx* ctx) { delete ctx; }

int kzg_commit_async(kzg_ctx* ctx, const kzg_srs* srs, const void* d_evals, size_t n, int window_bits, void* d_out,
                     int* d_err) {
    if (!ctx || !srs || !d_evals || !d_out || !d_err) return KZ_CRYPTO_ERR_INPUT;
    if (n == 0 || n > srs->n || log2_exact(n) < 0 || log2_exact(n) > (srs->curve == 2 ? 32 : 28))
        return KZ_CRYPTO_ERR_INPUT;
    if (window_bits != 0 && (window_bits < 5 || window_bits > 18)) return KZ_CRYPTO_ERR_INPUT;
    try {
        commit_device(ctx, srs, static_cast<const uint32_t*>(d_evals), n, window_bits, static_cast<uint8_t*>(d_out),
                      d_err);
        return KZ_CRYPTO_OK;
    } catch (const std::exception& ex) {
        set_last_error(ex.what());
        return KZ_CRYPTO_ERR_BACKEND;
    }
}

}  // extern "C"
