// Scalar-field NTT over BLS12-381 Fr and BN254 Fr on gfx950.
//
// The reference has no working Fr NTT (its ntt/ module is 64-bit-word only
// and the Metal Fr FFT in kzg/gpu/metal/kzg.metal:379-435 is never
// dispatched); the transform it describes, and that the CPU oracle restates
// after ntt/cpp/ntt.cpp:97-172, is the plain DFT X_k = sum_j x_j w^(jk) in
// natural order, inverse scaled by n^-1.
//
// Schedule: n = A * B * C (each <= 2^10), one to three passes of the same
// kernel.  A pass runs independent sub-DFTs of length L <= 1024 held in LDS
// (DIT: bit-reversed load, log2 L radix-2 stages fused in pairs into radix-4
// register passes, twiddles w_L^k read from the resident table), optionally
// multiplies its outputs by the four-step twiddle w_n^(e*k), and writes them
// back through a strided address map:
//   pass 1  columns of length A, stride B*C, twiddle w_n^(m k)      (in place)
//   pass 2  columns of length B inside each row, twiddle w_(BC)^(c k) (in place)
//   pass 3  contiguous rows of length C, written transposed to natural order
// Field elements stay canonical throughout: twiddles are stored in
// Montgomery form, so a Montgomery product x * (w R) * R^-1 = x w.  Twiddle
// tables w_n^e (e < n) live on the device per (field, n, direction).
#include <map>
#include <memory>
#include <mutex>
#include <tuple>

#include "engine.h"
#include "field.h"
#include "kinet/crypto/hip_runtime.h"
#include "ntt.h"

namespace kz {

constexpr int kNttTile = 1024;   // elements per workgroup tile (LDS)
constexpr int kNttThreads = 256;

struct NttPass {
    uint32_t logL, T, Q0;
    uint32_t in_b0, in_b1, in_s;
    uint32_t out_b0, out_b1, out_s;
    uint32_t tw_b0, tw_b1;   // four-step twiddle exponent base; both 0 = none
    uint32_t root_shift;     // w_L^k = w_n^(k << root_shift)
    uint32_t scale;          // multiply outputs by n^-1
    uint32_t check;          // flag inputs >= r
};

template <class F>
KZ_DEV Fe<F> fe_load_words9(const uint32_t* p) {
    Fe<F> r;
#pragma unroll
    for (int i = 0; i < F::N; ++i) r.v[i] = p[i];
    return r;
}

template <class F>
KZ_DEV void fe_store_words9(uint32_t* p, const Fe<F>& a) {
#pragma unroll
    for (int i = 0; i < F::N; ++i) p[i] = a.v[i];
}

__device__ __forceinline__ uint32_t bitrev(uint32_t j, uint32_t bits) {
    return bits ? (__brev(j) >> (32 - bits)) : 0u;
}

template <class F>
__global__ void __launch_bounds__(kNttThreads)
k_ntt_pass(const uint32_t* __restrict__ in, uint32_t* __restrict__ out, const uint32_t* __restrict__ tw,
           const uint32_t* __restrict__ ninv, NttPass d, uint32_t n_mask, int* __restrict__ err) {
    static_assert(F::N == 9, "Fr limbs");
    __shared__ uint32_t sh[kNttTile * F::N];
    // blockIdx.y selects one transform of a batch laid back to back (n words x 8)
    const size_t xform = (size_t)blockIdx.y * ((size_t)n_mask + 1) * 8;
    in += xform;
    out += xform;
    const uint32_t L = 1u << d.logL, tile = d.T * L;
    // Workgroup ids go round-robin over the 8 XCDs (each with its own L2).
    // Hand each XCD a contiguous range of tiles instead, so the neighbouring
    // columns of a strided pass -- which share 128-byte lines -- are read
    // through one L2.
    const uint32_t ng = gridDim.x, bid = blockIdx.x;
    const uint32_t tile_id = (ng % 8 == 0) ? (bid % 8) * (ng / 8) + bid / 8 : bid;
    const uint32_t q_start = tile_id * d.T;
    const uint32_t tid = threadIdx.x;

    // twiddles w_L^k = w_n^(k << root_shift) come straight from the resident
    // table (a few hundred lines, hot in L2): LDS holds only the tile, so four
    // workgroups fit on a CU
    const uint32_t tws = d.root_shift;

    bool bad = false;
    for (uint32_t e = tid; e < tile; e += kNttThreads) {
        uint32_t b, j;
        if (d.in_s == 1) { b = e >> d.logL; j = e & (L - 1); }
        else { b = e % d.T; j = e / d.T; }
        const uint32_t q = q_start + b, q0 = q % d.Q0, q1 = q / d.Q0;
        const size_t idx = (size_t)q0 * d.in_b0 + (size_t)q1 * d.in_b1 + (size_t)j * d.in_s;
        const uint4* src = reinterpret_cast<const uint4*>(in + idx * 8);
        const uint4 lo = src[0], hi = src[1];
        const uint32_t w[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
        const Fe<F> x = fe_from_words<F>(w, 8);
        if (d.check && !fe_is_canonical<F>(x)) bad = true;
        fe_store_words9<F>(sh + (b * L + bitrev(j, d.logL)) * F::N, x);
    }
    if (bad) atomicOr(err, 1);
    __syncthreads();

    // Butterflies: radix-2 DIT stages s = 0 .. logL-1 (stage s pairs elements
    // h = 2^s apart with twiddle w_(2h)^i = w_L^(i << (logL-1-s))), fused two
    // at a time: a group of 4h elements holds, for each i < h, the four
    // elements e, e+h, e+2h, e+3h that stages s and s+1 combine among
    // themselves, so they stay in registers across both stages and LDS is
    // read and written once per two stages.  An odd logL starts with the
    // twiddle-free stage 0 alone.  Groups never straddle sub-DFTs (4h | L).
    uint32_t s = 0;
    if (d.logL & 1) {
        for (uint32_t p = tid; p < tile / 2; p += kNttThreads) {
            const Fe<F> u = fe_load_words9<F>(sh + (2 * p) * F::N);
            const Fe<F> v = fe_load_words9<F>(sh + (2 * p + 1) * F::N);
            fe_store_words9<F>(sh + (2 * p) * F::N, fe_add<F>(u, v));
            fe_store_words9<F>(sh + (2 * p + 1) * F::N, fe_sub<F>(u, v));
        }
        __syncthreads();
        s = 1;
    }
    for (; s < d.logL; s += 2) {
        const uint32_t h = 1u << s, sh1 = d.logL - 1 - s, sh2 = d.logL - 2 - s;
        for (uint32_t q = tid; q < tile / 4; q += kNttThreads) {
            const uint32_t i = q & (h - 1), e = ((q >> s) << (s + 2)) + i;
            Fe<F> x0 = fe_load_words9<F>(sh + e * F::N);
            Fe<F> x1 = fe_load_words9<F>(sh + (e + h) * F::N);
            Fe<F> x2 = fe_load_words9<F>(sh + (e + 2 * h) * F::N);
            Fe<F> x3 = fe_load_words9<F>(sh + (e + 3 * h) * F::N);
            if (s != 0) {  // stage s: w_(2h)^i (= 1 at s = 0)
                const Fe<F> w = fe_load_words9<F>(tw + ((size_t)(i << sh1) << tws) * F::N);
                x1 = fe_mul<F>(x1, w);
                x3 = fe_mul<F>(x3, w);
            }
            const Fe<F> a0 = fe_add<F>(x0, x1), a1 = fe_sub<F>(x0, x1);
            // stage s+1: pairs (e, e+2h) with w_(4h)^i, (e+h, e+3h) with w_(4h)^(i+h)
            const Fe<F> a2 = fe_mul<F>(fe_add<F>(x2, x3), fe_load_words9<F>(tw + ((size_t)(i << sh2) << tws) * F::N));
            const Fe<F> a3 =
                fe_mul<F>(fe_sub<F>(x2, x3), fe_load_words9<F>(tw + ((size_t)((i + h) << sh2) << tws) * F::N));
            fe_store_words9<F>(sh + e * F::N, fe_add<F>(a0, a2));
            fe_store_words9<F>(sh + (e + 2 * h) * F::N, fe_sub<F>(a0, a2));
            fe_store_words9<F>(sh + (e + h) * F::N, fe_add<F>(a1, a3));
            fe_store_words9<F>(sh + (e + 3 * h) * F::N, fe_sub<F>(a1, a3));
        }
        __syncthreads();
    }

    const Fe<F> scale = d.scale ? fe_load_words9<F>(ninv) : fe_zero<F>();
    for (uint32_t e = tid; e < tile; e += kNttThreads) {
        uint32_t b, k;
        if (d.out_s == 1) { b = e >> d.logL; k = e & (L - 1); }
        else { b = e % d.T; k = e / d.T; }
        const uint32_t q = q_start + b, q0 = q % d.Q0, q1 = q / d.Q0;
        Fe<F> x = fe_load_words9<F>(sh + (b * L + k) * F::N);
        if (d.tw_b0 | d.tw_b1) {
            const uint32_t base = (q0 * d.tw_b0 + q1 * d.tw_b1) & n_mask;
            const uint32_t ex = (uint32_t)(((uint64_t)base * k) & n_mask);
            if (ex) x = fe_mul<F>(x, fe_load_words9<F>(tw + (size_t)ex * F::N));
        }
        if (d.scale) x = fe_mul<F>(x, scale);
        uint32_t w[8];
        fe_to_words<F>(x, w, 8);
        const size_t idx = (size_t)q0 * d.out_b0 + (size_t)q1 * d.out_b1 + (size_t)k * d.out_s;
        uint4* dst = reinterpret_cast<uint4*>(out + idx * 8);
        dst[0] = make_uint4(w[0], w[1], w[2], w[3]);
        dst[1] = make_uint4(w[4], w[5], w[6], w[7]);
    }
}

// ------------------------------------------------------------ twiddles
// root = g^exp (exp = (r-1) >> logn, 8 LE words), inverted for the inverse
// transform; ninv = n^-1.  Both in Montgomery form.
template <class F>
__global__ void k_ntt_root(const uint32_t* __restrict__ exp_words, uint32_t g, uint32_t logn, uint32_t inverse,
                           uint32_t* __restrict__ root_out, uint32_t* __restrict__ ninv_out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    Fe<F> base = fe_zero<F>();
    base.v[0] = g;
    base = fe_to_mont<F>(base);
    Fe<F> acc = fe_one<F>();
    for (int bit = 255; bit >= 0; --bit) {
        acc = fe_mul<F>(acc, acc);
        if ((exp_words[bit >> 5] >> (bit & 31)) & 1u) acc = fe_mul<F>(acc, base);
    }
    if (inverse) acc = fe_inv<F>(acc);
    fe_store_words9<F>(root_out, acc);
    Fe<F> nn = fe_zero<F>();
    nn.v[logn / kLimbBits] = 1u << (logn % kLimbBits);
    fe_store_words9<F>(ninv_out, fe_inv<F>(fe_to_mont<F>(nn)));
}

// tw[e] = root^e for e < n, by square-and-multiply per entry.
template <class F>
__global__ void __launch_bounds__(256) k_ntt_table(const uint32_t* __restrict__ root, uint32_t n,
                                                   uint32_t* __restrict__ tw) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const Fe<F> w = fe_load_words9<F>(root);
    Fe<F> acc = fe_one<F>(), base = w;
    for (uint32_t x = e; x; x >>= 1) {
        if (x & 1u) acc = fe_mul<F>(acc, base);
        base = fe_mul<F>(base, base);
    }
    fe_store_words9<F>(tw + (size_t)e * F::N, acc);
}

namespace {

struct NttTables {
    DevBuf tw, root, ninv, expw;
};

std::mutex g_ntt_mu;
std::map<std::tuple<int, int, int>, std::unique_ptr<NttTables>> g_ntt_tables;

// (r - 1) of each Fr as 4 little-endian 64-bit limbs.
const uint64_t kRm1[2][4] = {
    {0xFFFFFFFF00000000ull, 0x53BDA402FFFE5BFEull, 0x3339D80809A1D805ull, 0x73EDA753299D7D48ull},  // BLS12-381
    {0x43E1F593F0000000ull, 0x2833E84879B97091ull, 0xB85045B68181585Dull, 0x30644E72E131A029ull},  // BN254
};
const uint32_t kGen[2] = {7, 5};
const int kTwoAdicity[2] = {32, 28};

const NttTables& tables(int field, int logn, bool inverse, hipStream_t s) {
    std::lock_guard<std::mutex> g(g_ntt_mu);
    auto key = std::make_tuple(field, logn, inverse ? 1 : 0);
    auto it = g_ntt_tables.find(key);
    if (it != g_ntt_tables.end()) return *it->second;
    auto t = std::make_unique<NttTables>();
    // exponent (r - 1) >> logn as 8 words
    uint64_t e[4];
    for (int i = 0; i < 4; ++i) e[i] = kRm1[field][i];
    for (int k = 0; k < logn; ++k)
        for (int i = 0; i < 4; ++i) e[i] = (e[i] >> 1) | (i < 3 ? e[i + 1] << 63 : 0);
    uint32_t ew[8];
    for (int i = 0; i < 4; ++i) {
        ew[2 * i] = (uint32_t)e[i];
        ew[2 * i + 1] = (uint32_t)(e[i] >> 32);
    }
    const size_t n = (size_t)1 << logn;
    t->expw.ensure(sizeof ew);
    t->root.ensure(9 * 4);
    t->ninv.ensure(9 * 4);
    t->tw.ensure(n * 9 * 4);
    KZ_HIP(hipMemcpyAsync(t->expw.ptr, ew, sizeof ew, hipMemcpyHostToDevice, s));
    if (field == 0) {
        hipLaunchKernelGGL(k_ntt_root<BlsFr>, dim3(1), dim3(64), 0, s, t->expw.as<uint32_t>(), kGen[0],
                           (uint32_t)logn, (uint32_t)inverse, t->root.as<uint32_t>(), t->ninv.as<uint32_t>());
        hipLaunchKernelGGL(k_ntt_table<BlsFr>, dim3(ceil_div(n, 256)), dim3(256), 0, s, t->root.as<uint32_t>(),
                           (uint32_t)n, t->tw.as<uint32_t>());
    } else {
        hipLaunchKernelGGL(k_ntt_root<BnFr>, dim3(1), dim3(64), 0, s, t->expw.as<uint32_t>(), kGen[1],
                           (uint32_t)logn, (uint32_t)inverse, t->root.as<uint32_t>(), t->ninv.as<uint32_t>());
        hipLaunchKernelGGL(k_ntt_table<BnFr>, dim3(ceil_div(n, 256)), dim3(256), 0, s, t->root.as<uint32_t>(),
                           (uint32_t)n, t->tw.as<uint32_t>());
    }
    KZ_HIP(hipGetLastError());
    KZ_HIP(hipStreamSynchronize(s));  // the exponent staging copy is from the stack
    auto& ref = *t;
    g_ntt_tables[key] = std::move(t);
    return ref;
}

// Tuned pass splits per (field, log2 n): log2 of the sub-transform lengths of
// passes 1..3 (unused passes 0); all zero = the built-in split.
std::mutex g_sched_mu;
int g_sched[2][kNttMaxLog + 1][3];

void builtin_split(int logn, int (&p)[3]) {
    p[0] = p[1] = p[2] = 0;
    if (logn <= 10) p[0] = logn;
    else if (logn <= 20) { p[0] = (logn + 1) / 2; p[1] = logn - p[0]; }
    else { p[0] = (logn + 2) / 3; p[1] = (logn - p[0] + 1) / 2; p[2] = logn - p[0] - p[1]; }
}

void plan(int field, int logn, bool inverse, NttPass* passes, int& np) {
    int sp[3];
    ntt_schedule_of(field, logn, sp[0], sp[1], sp[2]);
    // la, lb, lc: lengths of the column pass, the inner column pass and the
    // row pass (la = lb = 0 for one pass, lb = 0 for two)
    int la, lb, lc;
    if (sp[2]) { la = sp[0]; lb = sp[1]; lc = sp[2]; np = 3; }
    else if (sp[1]) { la = sp[0]; lb = 0; lc = sp[1]; np = 2; }
    else { la = 0; lb = 0; lc = sp[0]; np = 1; }
    const uint32_t A = 1u << la, B = 1u << lb, C = 1u << lc, M = B * C;
    // sub-DFTs per workgroup: fill the 1024-element tile, but never more than
    // the n / L sub-DFTs the pass has
    auto T = [logn](int logL) {
        const uint32_t fit = (uint32_t)kNttTile >> logL, have = 1u << (logn - logL);
        return fit < have ? fit : have;
    };
    int k = 0;
    if (np >= 2) {
        NttPass p{};
        p.logL = la; p.T = T(la); p.Q0 = M;
        p.in_b0 = 1; p.in_b1 = 0; p.in_s = M;
        p.out_b0 = 1; p.out_b1 = 0; p.out_s = M;
        p.tw_b0 = 1; p.tw_b1 = 0; p.root_shift = lb + lc;
        passes[k++] = p;
    }
    if (np == 3) {
        NttPass p{};
        p.logL = lb; p.T = T(lb); p.Q0 = C;
        p.in_b0 = 1; p.in_b1 = M; p.in_s = C;
        p.out_b0 = 1; p.out_b1 = M; p.out_s = C;
        p.tw_b0 = A; p.tw_b1 = 0; p.root_shift = la + lc;
        passes[k++] = p;
    }
    NttPass p{};
    p.logL = lc; p.T = T(lc); p.Q0 = A;
    p.in_b0 = M; p.in_b1 = C; p.in_s = 1;
    p.out_b0 = 1; p.out_b1 = A; p.out_s = A * B;
    p.root_shift = la + lb;
    passes[k++] = p;
    passes[0].check = 1;
    passes[np - 1].scale = inverse ? 1 : 0;
}

}  // namespace

bool ntt_split_valid(int logn, int p1, int p2, int p3) {
    if (logn < 0 || logn > kNttMaxLog) return false;
    if (p1 < 0 || p2 < 0 || p3 < 0 || p1 > 10 || p2 > 10 || p3 > 10) return false;
    if (p1 + p2 + p3 != logn) return false;
    if (p3 && (!p1 || !p2)) return false;  // zeros only trailing
    if (p2 && !p1) return false;
    return true;
}

void ntt_set_schedule(int field, int logn, int p1, int p2, int p3) {
    if (field < 0 || field > 1 || logn < 0 || logn > kNttMaxLog) throw std::invalid_argument("NTT size class");
    if (!(p1 == 0 && p2 == 0 && p3 == 0) && !ntt_split_valid(logn, p1, p2, p3))
        throw std::invalid_argument("NTT pass split must sum to log2 n with passes of at most 2^10");
    std::lock_guard<std::mutex> g(g_sched_mu);
    g_sched[field][logn][0] = p1;
    g_sched[field][logn][1] = p2;
    g_sched[field][logn][2] = p3;
}

void ntt_get_schedule(int field, int logn, int& p1, int& p2, int& p3) {
    if (field < 0 || field > 1 || logn < 0 || logn > kNttMaxLog) throw std::invalid_argument("NTT size class");
    std::lock_guard<std::mutex> g(g_sched_mu);
    p1 = g_sched[field][logn][0];
    p2 = g_sched[field][logn][1];
    p3 = g_sched[field][logn][2];
}

void ntt_schedule_of(int field, int logn, int& p1, int& p2, int& p3) {
    ntt_get_schedule(field, logn, p1, p2, p3);
    if (p1 | p2 | p3) return;
    int b[3];
    builtin_split(logn, b);
    p1 = b[0];
    p2 = b[1];
    p3 = b[2];
}

void ntt_prepare(int field, int logn, bool inverse, hipStream_t s) {
    if (field < 0 || field > 1 || logn < 0 || logn > kNttMaxLog || logn > kTwoAdicity[field])
        throw std::invalid_argument("NTT size out of range");
    (void)tables(field, logn, inverse, s);
}

void ntt_device_batch(int field, int logn, bool inverse, const uint32_t* d_in, uint32_t* d_out, uint32_t* d_tmp,
                      size_t count, int* d_err, hipStream_t s) {
    if (field < 0 || field > 1 || logn < 0 || logn > kNttMaxLog || logn > kTwoAdicity[field])
        throw std::invalid_argument("NTT size out of range");
    if (count == 0) return;
    if (count > 65535) throw std::invalid_argument("NTT batch larger than 65535 transforms");
    const NttTables& t = tables(field, logn, inverse, s);
    NttPass passes[3];
    int np = 0;
    plan(field, logn, inverse, passes, np);
    const uint32_t n = 1u << logn;
    for (int i = 0; i < np; ++i) {
        const uint32_t* src = (i == 0) ? d_in : d_tmp;
        uint32_t* dst = (i == np - 1) ? d_out : d_tmp;
        const unsigned grid = (n >> passes[i].logL) / passes[i].T;  // exact: powers of two, T <= n / L
        if (field == 0)
            hipLaunchKernelGGL(k_ntt_pass<BlsFr>, dim3(grid, (unsigned)count), dim3(kNttThreads), 0, s, src, dst,
                               t.tw.as<uint32_t>(), t.ninv.as<uint32_t>(), passes[i], n - 1, d_err);
        else
            hipLaunchKernelGGL(k_ntt_pass<BnFr>, dim3(grid, (unsigned)count), dim3(kNttThreads), 0, s, src, dst,
                               t.tw.as<uint32_t>(), t.ninv.as<uint32_t>(), passes[i], n - 1, d_err);
    }
    KZ_HIP(hipGetLastError());
}

void ntt_device(int field, int logn, bool inverse, const uint32_t* d_in, uint32_t* d_out, uint32_t* d_tmp,
                int* d_err, hipStream_t s) {
    ntt_device_batch(field, logn, inverse, d_in, d_out, d_tmp, 1, d_err, s);
}

}  // namespace kz

using namespace kz;

extern "C" {

int kinet_hip_ntt_fr(uint32_t field, int inverse, const void* d_in, void* d_out, void* d_tmp, size_t n, int* d_err,
                     void* stream) {
    if (field > 1) return KZ_CRYPTO_ERR_INPUT;
    if (!d_in || !d_out || !d_tmp || !d_err) return KZ_CRYPTO_ERR_INPUT;
    int logn = 0;
    while (((size_t)1 << logn) < n) ++logn;
    if (n == 0 || ((size_t)1 << logn) != n || logn > kNttMaxLog || logn > (field == 0 ? 32 : 28))
        return KZ_CRYPTO_ERR_INPUT;
    try {
        (void)engine();
        ntt_device((int)field, logn, inverse != 0, static_cast<const uint32_t*>(d_in), static_cast<uint32_t*>(d_out),
                   static_cast<uint32_t*>(d_tmp), d_err, static_cast<hipStream_t>(stream));
        return KZ_CRYPTO_OK;
    } catch (const std::exception& ex) {
        set_last_error(ex.what());
        return KZ_CRYPTO_ERR_BACKEND;
    }
}

int kinet_hip_ntt_fr_batch(uint32_t field, int inverse, const void* d_in, void* d_out, void* d_tmp, size_t n,
                           size_t count, int* d_err, void* stream) {
    if (field > 1) return KZ_CRYPTO_ERR_INPUT;
    if (!d_in || !d_out || !d_tmp || !d_err) return KZ_CRYPTO_ERR_INPUT;
    int logn = 0;
    while (((size_t)1 << logn) < n) ++logn;
    if (n == 0 || ((size_t)1 << logn) != n || logn > kNttMaxLog || logn > (field == 0 ? 32 : 28))
        return KZ_CRYPTO_ERR_INPUT;
    if (count > 65535) return KZ_CRYPTO_ERR_INPUT;
    try {
        (void)engine();
        ntt_device_batch((int)field, logn, inverse != 0, static_cast<const uint32_t*>(d_in),
                         static_cast<uint32_t*>(d_out), static_cast<uint32_t*>(d_tmp), count, d_err,
                         static_cast<hipStream_t>(stream));
        return KZ_CRYPTO_OK;
    } catch (const std::exception& ex) {
        set_last_error(ex.what());
        return KZ_CRYPTO_ERR_BACKEND;
    }
}

static int ntt_host(uint32_t field, uint8_t* elems, size_t n, bool inverse) {
    if (!elems || field > 1) return KZ_CRYPTO_ERR_INPUT;
    int logn = 0;
    while (((size_t)1 << logn) < n) ++logn;
    if (n == 0 || ((size_t)1 << logn) != n || logn > kNttMaxLog || logn > (field == 0 ? 32 : 28))
        return KZ_CRYPTO_ERR_INPUT;
    try {
        Engine& e = engine();
        std::lock_guard<std::mutex> g(e.mu);
        hipStream_t s = e.stream;
        uint32_t* d_a = static_cast<uint32_t*>(e.ntt_a.ensure(32 * n));
        uint32_t* d_b = static_cast<uint32_t*>(e.ntt_b.ensure(32 * n));
        int* d_err = static_cast<int*>(e.ntt_tw.ensure(sizeof(int)));
        KZ_HIP(hipMemsetAsync(d_err, 0, sizeof(int), s));
        KZ_HIP(hipMemcpyAsync(d_a, elems, 32 * n, hipMemcpyHostToDevice, s));
        ntt_device((int)field, logn, inverse, d_a, d_a, d_b, d_err, s);
        int err = 0;
        KZ_HIP(hipMemcpyAsync(&err, d_err, sizeof(int), hipMemcpyDeviceToHost, s));
        KZ_HIP(hipStreamSynchronize(s));
        if (err) return KZ_CRYPTO_ERR_INPUT;  // elems untouched
        KZ_HIP(hipMemcpyAsync(elems, d_a, 32 * n, hipMemcpyDeviceToHost, s));
        KZ_HIP(hipStreamSynchronize(s));
        return KZ_CRYPTO_OK;
    } catch (const std::exception& ex) {
        set_last_error(ex.what());
        return KZ_CRYPTO_ERR_BACKEND;
    }
}

int ntt_fr_forward(uint32_t field, uint8_t* elems, size_t n) { return ntt_host(field, elems, n, false); }
int ntt_fr_inverse(uint32_t field, uint8_t* elems, size_t n) { return ntt_host(field, elems, n, true); }

}  // extern "C"
