// EIP-4844 blob entries on the device: kzg_blob_to_commit and
// kzg_commit_to_proof (c-abi/crypto.h:238-239), the operations the reference
// hands to c-kzg-4844 v2.1.7 (kzg/cpp/kzg_blob.cpp:69-81, 105-125), and the
// trusted-setup loader they need (the reference embeds the mainnet
// trusted_setup.txt at configure time and loads it once, kzg_blob.cpp:28-47).
//
// The algorithm is EIP-4844's (consensus-specs deneb polynomial-commitments):
//   blob            4096 field elements, 32-byte big-endian, each < r
//   commitment      g1_lincomb(bit_reversal_permutation(SETUP_G1_LAGRANGE), blob)
//   proof at z      y = f(z) by the barycentric formula over the domain
//                   w_i = w^brp(i) (or f_i when z = w_i), quotient evaluations
//                   q_i = (f_i - y) / (w_i - z), except at z = w_m:
//                   q_m = sum_{i != m} (f_i - y) w_i / (z (z - w_i)); then
//                   proof = g1_lincomb(bit_reversal_permutation(LAGRANGE), q)
// with w the 4096-th root of unity 7^((r - 1) / 4096).  The bit-reversed
// Lagrange points are one SRS handle, so both are one MSM of 4096 points.
//
// Setup file: c-kzg-4844's text format -- "4096", the G2 point count, 4096
// compressed G1 points in Lagrange form (natural order, 96 hex digits each),
// the G2 points (192 hex digits each), then, in files of c-kzg 2.x, 4096
// compressed G1 points in monomial form.  G1 points are decompressed and
// checked on the curve and in G1 on the device; G2 points are read for
// their count and encoding only (they serve verification, which this library
// does not provide).
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "engine.h"
#include "field.h"
#include "kinet/crypto/hip_runtime.h"
#include "kinet/crypto/kzg.h"
#include "kzg_internal.h"
#include "msm.h"

namespace kz {
bool canonical_fr(int curve, const uint8_t* le32);

namespace {

constexpr size_t kBlobN = 4096;  // FIELD_ELEMENTS_PER_BLOB
constexpr int kBlobLog = 12;
using Fr = BlsFr;

std::mutex g_setup_mu;
kzg_srs* g_setup = nullptr;  // bit-reversed Lagrange points, or null

uint32_t brp12(uint32_t i) { return __builtin_bitreverse32(i) >> (32 - kBlobLog); }

// ------------------------------------------------------------- parsing
int hexval(char c) {
    if (c >= '0' && c <= '9') return c - '0';
    if (c >= 'a' && c <= 'f') return c - 'a' + 10;
    if (c >= 'A' && c <= 'F') return c - 'A' + 10;
    return -1;
}

struct Tokens {
    const std::string& s;
    size_t pos = 0;
    explicit Tokens(const std::string& text) : s(text) {}
    bool next(std::string& tok) {
        while (pos < s.size() && std::isspace((unsigned char)s[pos])) ++pos;
        if (pos >= s.size()) return false;
        const size_t b = pos;
        while (pos < s.size() && !std::isspace((unsigned char)s[pos])) ++pos;
        tok.assign(s, b, pos - b);
        return true;
    }
};

bool parse_hex(const std::string& tok, size_t nbytes, uint8_t* out) {
    if (tok.size() != 2 * nbytes) return false;
    for (size_t k = 0; k < nbytes; ++k) {
        const int hi = hexval(tok[2 * k]), lo = hexval(tok[2 * k + 1]);
        if (hi < 0 || lo < 0) return false;
        out[k] = (uint8_t)(hi << 4 | lo);
    }
    return true;
}

bool parse_count(const std::string& tok, size_t& v) {
    if (tok.empty() || tok.size() > 9) return false;
    v = 0;
    for (char c : tok) {
        if (c < '0' || c > '9') return false;
        v = v * 10 + (size_t)(c - '0');
    }
    return true;
}

// Parsed setup: compressed G1 Lagrange points (natural order) and, if the
// file has them, the G1 monomial points.  Returns an empty string on success,
// else the reason.
std::string parse_setup(const std::string& text, std::vector<uint8_t>& lag, std::vector<uint8_t>& mono) {
    Tokens t(text);
    std::string tok;
    size_t n1 = 0, n2 = 0;
    if (!t.next(tok) || !parse_count(tok, n1)) return "missing G1 point count";
    if (!t.next(tok) || !parse_count(tok, n2)) return "missing G2 point count";
    if (n1 != kBlobN) return "the G1 point count must be 4096 (FIELD_ELEMENTS_PER_BLOB)";
    if (n2 < 2 || n2 > 4096) return "the G2 point count must be at least 2";
    lag.resize(48 * n1);
    for (size_t i = 0; i < n1; ++i)
        if (!t.next(tok) || !parse_hex(tok, 48, &lag[48 * i])) return "bad G1 Lagrange point " + std::to_string(i);
    uint8_t g2[96];
    for (size_t i = 0; i < n2; ++i) {
        if (!t.next(tok) || !parse_hex(tok, 96, g2)) return "bad G2 point " + std::to_string(i);
        if (!(g2[0] & 0x80)) return "G2 point " + std::to_string(i) + " is not in compressed form";
    }
    mono.clear();
    if (t.next(tok)) {
        mono.resize(48 * n1);
        if (!parse_hex(tok, 48, &mono[0])) return "bad G1 monomial point 0";
        for (size_t i = 1; i < n1; ++i)
            if (!t.next(tok) || !parse_hex(tok, 48, &mono[48 * i])) return "bad G1 monomial point " + std::to_string(i);
        if (t.next(tok)) return "trailing data after the G1 monomial points";
    }
    return "";
}

// Decompresses n 48-byte G1 points on the device into host wire records;
// false if any encoding is rejected.
bool decompress_g1(const uint8_t* comp, size_t n, std::vector<uint8_t>& wire) {
    Engine& e = engine();
    std::lock_guard<std::mutex> g(e.mu);
    hipStream_t s = e.stream;
    uint8_t* d_comp = static_cast<uint8_t*>(e.h2d_a.ensure(48 * n));
    uint8_t* d_wire = static_cast<uint8_t*>(e.h2d_b.ensure(96 * n));
    int* d_bad = static_cast<int*>(e.out.ensure(16));
    KZ_HIP(hipMemcpyAsync(d_comp, comp, 48 * n, hipMemcpyHostToDevice, s));
    KZ_HIP(hipMemsetAsync(d_bad, 0, sizeof(int), s));
    if (kinet_hip_points_decompress(2, d_comp, n, d_wire, d_bad, s) != 0)
        throw std::runtime_error(std::string("decompression failed: ") + kinet_hip_last_error());
    wire.resize(96 * n);
    int bad = 1;
    KZ_HIP(hipMemcpyAsync(wire.data(), d_wire, 96 * n, hipMemcpyDeviceToHost, s));
    KZ_HIP(hipMemcpyAsync(&bad, d_bad, sizeof bad, hipMemcpyDeviceToHost, s));
    KZ_HIP(hipStreamSynchronize(s));
    return bad == 0;
}

// Loads the file named by KINET_KZG_TRUSTED_SETUP when no setup is loaded.
// Called with g_setup_mu held.
int ensure_setup_locked();

int load_setup_text_locked(const std::string& text) {
    std::vector<uint8_t> lag, mono;
    const std::string why = parse_setup(text, lag, mono);
    if (!why.empty()) {
        set_last_error("trusted setup: " + why);
        return KZ_CRYPTO_ERR_INPUT;
    }
    std::vector<uint8_t> wire, wire_mono;
    if (!decompress_g1(lag.data(), kBlobN, wire)) {
        set_last_error("trusted setup: a G1 Lagrange point is not a valid compressed curve point");
        return KZ_CRYPTO_ERR_INPUT;
    }
    if (!mono.empty() && !decompress_g1(mono.data(), kBlobN, wire_mono)) {
        set_last_error("trusted setup: a G1 monomial point is not a valid compressed curve point");
        return KZ_CRYPTO_ERR_INPUT;
    }
    // bit_reversal_permutation(KZG_SETUP_G1_LAGRANGE)
    std::vector<uint8_t> brp(96 * kBlobN);
    for (uint32_t i = 0; i < kBlobN; ++i) std::memcpy(&brp[96 * i], &wire[96 * brp12(i)], 96);
    kzg_srs* srs = nullptr;
    int rc = kzg_srs_create(2, brp.data(), kBlobN, &srs);
    if (rc != KZ_CRYPTO_OK) return rc;
    if (!srs->g1) {  // kzg_srs_create checked membership in G1
        kzg_srs_destroy(srs);
        set_last_error("trusted setup: a G1 Lagrange point lies outside G1");
        return KZ_CRYPTO_ERR_INPUT;
    }
    if (!wire_mono.empty()) {
        kzg_srs* m = nullptr;
        rc = kzg_srs_create(2, wire_mono.data(), kBlobN, &m);
        const bool ok = rc == KZ_CRYPTO_OK && m->g1;
        if (m) kzg_srs_destroy(m);
        if (!ok) {
            kzg_srs_destroy(srs);
            set_last_error("trusted setup: a G1 monomial point lies outside G1");
            return rc == KZ_CRYPTO_OK ? KZ_CRYPTO_ERR_INPUT : rc;
        }
    }
    if (g_setup) kzg_srs_destroy(g_setup);
    g_setup = srs;
    return KZ_CRYPTO_OK;
}

int load_setup_file_locked(const char* path) {
    FILE* f = std::fopen(path, "rb");
    if (!f) {
        set_last_error(std::string("trusted setup: cannot open ") + path);
        return KZ_CRYPTO_ERR_INPUT;
    }
    std::string text;
    char buf[1 << 16];
    size_t got;
    while ((got = std::fread(buf, 1, sizeof buf, f)) > 0) text.append(buf, got);
    std::fclose(f);
    return load_setup_text_locked(text);
}

int ensure_setup_locked() {
    if (g_setup) return KZ_CRYPTO_OK;
    const char* env = std::getenv("KINET_KZG_TRUSTED_SETUP");
    if (!env || !*env) {
        set_last_error("no trusted setup loaded (kzg_load_trusted_setup_file, or KINET_KZG_TRUSTED_SETUP)");
        return KZ_CRYPTO_ERR_BACKEND;
    }
    return load_setup_file_locked(env);
}

// 32-byte big-endian field elements -> little-endian records; false if any is >= r
bool blob_to_le(const uint8_t* be, size_t n, uint8_t* le) {
    for (size_t i = 0; i < n; ++i) {
        for (int k = 0; k < 32; ++k) le[32 * i + k] = be[32 * i + 31 - k];
        if (!canonical_fr(2, le + 32 * i)) return false;
    }
    return true;
}

// ------------------------------------------------------------- the proof
// Per i: w_i = w^brp(i) (Montgomery words), d_i = z - w_i; inv_i = 1 / d_i
// (0 where d_i = 0, whose index goes to *dom as i + 1); t_i = f_i w_i inv_i.
__global__ void __launch_bounds__(128)
k_blob_terms(const uint32_t* __restrict__ f_le, const uint32_t* __restrict__ z_le, uint32_t* __restrict__ w_out,
             uint32_t* __restrict__ inv_out, uint32_t* __restrict__ t_out, uint32_t* __restrict__ dom) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kBlobN) return;
    // w = 7^((r - 1) / 4096)
    uint32_t e[8];
    fe_to_words<Fr>(fe_const<Fr>(Fr::P), e, 8);
    e[0] -= 1;
    for (int k = 0; k < kBlobLog; ++k) {
        for (int j = 0; j < 7; ++j) e[j] = (e[j] >> 1) | (e[j + 1] << 31);
        e[7] >>= 1;
    }
    Fe<Fr> seven = fe_zero<Fr>();
    seven.v[0] = 7;
    const Fe<Fr> g = fe_to_mont<Fr>(seven);
    Fe<Fr> w = fe_one<Fr>();
#pragma unroll 1
    for (int bit = 255; bit >= 0; --bit) {
        w = fe_mul<Fr>(w, w);
        if ((e[bit >> 5] >> (bit & 31)) & 1u) w = fe_mul<Fr>(w, g);
    }
    const uint32_t ex = __brev(i) >> (32 - kBlobLog);
    Fe<Fr> wi = fe_one<Fr>(), base = w;
    for (uint32_t k = ex; k; k >>= 1) {
        if (k & 1u) wi = fe_mul<Fr>(wi, base);
        base = fe_mul<Fr>(base, base);
    }
    const Fe<Fr> z = fe_to_mont<Fr>(fe_from_words<Fr>(z_le, 8));
    const Fe<Fr> d = fe_sub<Fr>(z, wi);
    Fe<Fr> inv = fe_zero<Fr>(), t = fe_zero<Fr>();
    if (fe_is_zero<Fr>(d)) {
        atomicMax(dom, i + 1);
    } else {
        inv = fe_inv<Fr>(d);
        const Fe<Fr> f = fe_to_mont<Fr>(fe_from_words<Fr>(f_le + (size_t)i * 8, 8));
        t = fe_mul<Fr>(fe_mul<Fr>(f, wi), inv);
    }
#pragma unroll
    for (int k = 0; k < Fr::N; ++k) {
        w_out[(size_t)i * Fr::N + k] = wi.v[k];
        inv_out[(size_t)i * Fr::N + k] = inv.v[k];
        t_out[(size_t)i * Fr::N + k] = t.v[k];
    }
}

// out[b] = sum of in[j] over block b's grid-stride share (Montgomery words)
__global__ void __launch_bounds__(256) k_blob_sum(const uint32_t* __restrict__ in, uint32_t n,
                                                  uint32_t* __restrict__ out) {
    __shared__ uint32_t sh[256 * Fr::N];
    Fe<Fr> acc = fe_zero<Fr>();
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        Fe<Fr> v;
#pragma unroll
        for (int k = 0; k < Fr::N; ++k) v.v[k] = in[(size_t)j * Fr::N + k];
        acc = fe_add<Fr>(acc, v);
    }
#pragma unroll
    for (int k = 0; k < Fr::N; ++k) sh[threadIdx.x * Fr::N + k] = acc.v[k];
    __syncthreads();
    for (uint32_t st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) {
            Fe<Fr> a, b;
#pragma unroll
            for (int k = 0; k < Fr::N; ++k) {
                a.v[k] = sh[threadIdx.x * Fr::N + k];
                b.v[k] = sh[(threadIdx.x + st) * Fr::N + k];
            }
            a = fe_add<Fr>(a, b);
#pragma unroll
            for (int k = 0; k < Fr::N; ++k) sh[threadIdx.x * Fr::N + k] = a.v[k];
        }
        __syncthreads();
    }
    if (threadIdx.x < Fr::N) out[(size_t)blockIdx.x * Fr::N + threadIdx.x] = sh[threadIdx.x];
}

// y = f_m if z = w_m, else (z^4096 - 1) / 4096 * S; Montgomery words at y_mont
// and canonical LE words at y_le
__global__ void k_blob_y(const uint32_t* __restrict__ S_words, const uint32_t* __restrict__ f_le,
                         const uint32_t* __restrict__ z_le, const uint32_t* __restrict__ dom,
                         uint32_t* __restrict__ y_mont, uint32_t* __restrict__ y_le) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    Fe<Fr> y;
    if (*dom) {
        y = fe_to_mont<Fr>(fe_from_words<Fr>(f_le + (size_t)(*dom - 1) * 8, 8));
    } else {
        Fe<Fr> S;
#pragma unroll
        for (int k = 0; k < Fr::N; ++k) S.v[k] = S_words[k];
        Fe<Fr> zn = fe_to_mont<Fr>(fe_from_words<Fr>(z_le, 8));
        for (int k = 0; k < kBlobLog; ++k) zn = fe_mul<Fr>(zn, zn);
        Fe<Fr> nn = fe_zero<Fr>();
        nn.v[0] = (uint32_t)kBlobN;
        y = fe_mul<Fr>(fe_mul<Fr>(fe_sub<Fr>(zn, fe_one<Fr>()), fe_inv<Fr>(fe_to_mont<Fr>(nn))), S);
    }
#pragma unroll
    for (int k = 0; k < Fr::N; ++k) y_mont[k] = y.v[k];
    fe_to_words<Fr>(fe_from_mont<Fr>(y), y_le, 8);
}

// q_i = (y - f_i) inv_i = (f_i - y) / (w_i - z) as canonical LE words, and
// u_i = (f_i - y) w_i inv_i (Montgomery; 0 at the domain index), whose sum
// over 1 / z is the quotient at a domain point z
__global__ void __launch_bounds__(256)
k_blob_quotient(const uint32_t* __restrict__ f_le, const uint32_t* __restrict__ w, const uint32_t* __restrict__ inv,
                const uint32_t* __restrict__ y_mont, uint32_t* __restrict__ q_out, uint32_t* __restrict__ u_out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kBlobN) return;
    Fe<Fr> y, iv, wi;
#pragma unroll
    for (int k = 0; k < Fr::N; ++k) {
        y.v[k] = y_mont[k];
        iv.v[k] = inv[(size_t)i * Fr::N + k];
        wi.v[k] = w[(size_t)i * Fr::N + k];
    }
    const Fe<Fr> f = fe_to_mont<Fr>(fe_from_words<Fr>(f_le + (size_t)i * 8, 8));
    const Fe<Fr> fy = fe_sub<Fr>(f, y);
    const Fe<Fr> u = fe_mul<Fr>(fe_mul<Fr>(fy, wi), iv);
    fe_to_words<Fr>(fe_from_mont<Fr>(fe_neg<Fr>(fe_mul<Fr>(fy, iv))), q_out + (size_t)i * 8, 8);
#pragma unroll
    for (int k = 0; k < Fr::N; ++k) u_out[(size_t)i * Fr::N + k] = u.v[k];
}

// At a domain point z = w_m: q_m = (sum_i u_i) / z, since
// u_i = (f_i - y) w_i / (z - w_i) and the term of q_m is u_i / z.
__global__ void k_blob_domain_fix(const uint32_t* __restrict__ U_words, const uint32_t* __restrict__ z_le,
                                  const uint32_t* __restrict__ dom, uint32_t* __restrict__ q_out) {
    if (threadIdx.x != 0 || blockIdx.x != 0 || *dom == 0) return;
    Fe<Fr> U;
#pragma unroll
    for (int k = 0; k < Fr::N; ++k) U.v[k] = U_words[k];
    const Fe<Fr> z = fe_to_mont<Fr>(fe_from_words<Fr>(z_le, 8));  // z = w_m != 0
    fe_to_words<Fr>(fe_from_mont<Fr>(fe_mul<Fr>(U, fe_inv<Fr>(z))), q_out + (size_t)(*dom - 1) * 8, 8);
}

}  // namespace
}  // namespace kz

using namespace kz;

extern "C" {

int kzg_load_trusted_setup_file(const char* path) {
    if (!path) return KZ_CRYPTO_ERR_INPUT;
    try {
        std::lock_guard<std::mutex> g(g_setup_mu);
        return load_setup_file_locked(path);
    } catch (const std::exception& ex) {
        set_last_error(ex.what());
        return KZ_CRYPTO_ERR_BACKEND;
    }
}

void kzg_free_trusted_setup(void) {
    std::lock_guard<std::mutex> g(g_setup_mu);
    if (g_setup) kzg_srs_destroy(g_setup);
    g_setup = nullptr;
}

int kzg_blob_to_commit(const uint8_t blob[131072], uint8_t commit[48]) {
    if (!blob || !commit) return KZ_CRYPTO_ERR_INPUT;
    std::vector<uint8_t> le(32 * kBlobN);
    if (!blob_to_le(blob, kBlobN, le.data())) return KZ_CRYPTO_ERR_INPUT;
    try {
        std::lock_guard<std::mutex> g(g_setup_mu);
        const int rc = ensure_setup_locked();
        if (rc != KZ_CRYPTO_OK) return rc;
        return kzg_commit_coeffs(g_setup, le.data(), kBlobN, commit);
    } catch (const std::exception& ex) {
        set_last_error(ex.what());
        return KZ_CRYPTO_ERR_BACKEND;
    }
}

int kzg_commit_to_proof(const uint8_t blob[131072], const uint8_t z[32], uint8_t proof[48], uint8_t y[32]) {
    if (!blob || !z || !proof || !y) return KZ_CRYPTO_ERR_INPUT;
    std::vector<uint8_t> le(32 * kBlobN);
    uint8_t zle[32];
    if (!blob_to_le(blob, kBlobN, le.data()) || !blob_to_le(z, 1, zle)) return KZ_CRYPTO_ERR_INPUT;
    try {
        std::lock_guard<std::mutex> sg(g_setup_mu);
        const int rc = ensure_setup_locked();
        if (rc != KZ_CRYPTO_OK) return rc;
        Engine& e = engine();
        std::lock_guard<std::mutex> g(e.mu);
        hipStream_t s = e.stream;
        constexpr size_t W = Fr::N;
        uint32_t* d_f = static_cast<uint32_t*>(e.h2d_a.ensure(32 * kBlobN));
        uint32_t* base = static_cast<uint32_t*>(e.ntt_a.ensure((4 * W * kBlobN + 8 * kBlobN + 2 * 257 * W + 64) * 4));
        uint32_t* d_w = base;
        uint32_t* d_inv = d_w + W * kBlobN;
        uint32_t* d_t = d_inv + W * kBlobN;
        uint32_t* d_u = d_t + W * kBlobN;
        uint32_t* d_q = d_u + W * kBlobN;  // 32-byte canonical records
        uint32_t* d_partS = d_q + 8 * kBlobN;
        uint32_t* d_partU = d_partS + 257 * W;
        uint32_t* d_misc = d_partU + 257 * W;  // y (W), y LE (8), z LE (8), domain index
        uint32_t *d_y = d_misc, *d_yle = d_misc + 16, *d_z = d_misc + 32, *d_dom = d_misc + 48;
        uint8_t* d_out = static_cast<uint8_t*>(e.out.ensure(msm_out_bytes(2) + 16));
        KZ_HIP(hipMemcpyAsync(d_f, le.data(), 32 * kBlobN, hipMemcpyHostToDevice, s));
        KZ_HIP(hipMemcpyAsync(d_z, zle, 32, hipMemcpyHostToDevice, s));
        KZ_HIP(hipMemsetAsync(d_dom, 0, sizeof(uint32_t), s));
        hipLaunchKernelGGL(k_blob_terms, dim3(kBlobN / 128), dim3(128), 0, s, d_f, d_z, d_w, d_inv, d_t, d_dom);
        hipLaunchKernelGGL(k_blob_sum, dim3(16), dim3(256), 0, s, d_t, (uint32_t)kBlobN, d_partS);
        hipLaunchKernelGGL(k_blob_sum, dim3(1), dim3(256), 0, s, d_partS, 16u, d_partS + 256 * W);
        hipLaunchKernelGGL(k_blob_y, dim3(1), dim3(64), 0, s, d_partS + 256 * W, d_f, d_z, d_dom, d_y, d_yle);
        hipLaunchKernelGGL(k_blob_quotient, dim3(kBlobN / 256), dim3(256), 0, s, d_f, d_w, d_inv, d_y, d_q, d_u);
        hipLaunchKernelGGL(k_blob_sum, dim3(16), dim3(256), 0, s, d_u, (uint32_t)kBlobN, d_partU);
        hipLaunchKernelGGL(k_blob_sum, dim3(1), dim3(256), 0, s, d_partU, 16u, d_partU + 256 * W);
        hipLaunchKernelGGL(k_blob_domain_fix, dim3(1), dim3(64), 0, s, d_partU + 256 * W, d_z, d_dom, d_q);
        KZ_HIP(hipGetLastError());
        srs_msm(g_setup, d_q, kBlobN, 0, e.ws, s, d_out);
        uint8_t host[96 + 48], yle[32];
        KZ_HIP(hipMemcpyAsync(host, d_out, msm_out_bytes(2), hipMemcpyDeviceToHost, s));
        KZ_HIP(hipMemcpyAsync(yle, d_yle, 32, hipMemcpyDeviceToHost, s));
        KZ_HIP(hipStreamSynchronize(s));
        std::memcpy(proof, host + 96, 48);
        for (int k = 0; k < 32; ++k) y[k] = yle[31 - k];  // big-endian
        return KZ_CRYPTO_OK;
    } catch (const std::exception& ex) {
        set_last_error(ex.what());
        return KZ_CRYPTO_ERR_BACKEND;
    }
}

}  // extern "C"
