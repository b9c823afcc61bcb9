// Scalar-field NTT over BLS12-381 Fr and BN254 Fr on gfx950.
//
// The reference has no working Fr NTT (its ntt/ module is 64-bit-word only
// and the Metal Fr FFT in kzg/gpu/metal/kzg.metal:379-435 is never
// dispatched); the transform it describes, and that the CPU oracle restates
// after ntt/cpp/ntt.cpp:97-172, is the plain DFT X_k = sum_j x_j w^(jk) in
// natural order, inverse scaled by n^-1.
//
// Schedule: n = A * B * C (each <= 2^11), one to three passes of the same
// kernel.  A pass runs independent sub-DFTs of length L <= 2048 held in LDS
// (DIT: bit-reversed load, log2 L radix-2 stages fused in pairs into radix-4
// register passes), multiplies its outputs by the four-step twiddle, and
// writes them back through a strided address map:
//   pass 1  columns of length A, stride B*C, twiddle w_n^(m k)      (in place)
//   pass 2  columns of length B inside each row, twiddle w_(BC)^(c k) (in place)
//   pass 3  contiguous rows of length C, written transposed to natural order
// Values are reduced lazily inside a pass (ntt_math.h): a pass's outputs are
// its four-step products (below 2r, stored as 8 words) or, in the last pass,
// canonical.  Twiddles are canonical Montgomery forms, so a Montgomery
// product x * (w R) * R^-1 = x w.  Twiddle tables per (field, direction):
//   sub-DFT tables  w_L^e for e < L/2, per L
//   four-step       one per (n, split) and pass, in the order the pass reads
//                   it: entry k * Q0 + q0 is w_n^(mult q0 k) for output k of
//                   column q0, so a workgroup's T neighbouring columns read T
//                   neighbouring entries.
#include <map>
#include <memory>
#include <mutex>
#include <tuple>

#include "engine.h"
#include "field.h"
#include "kinet/crypto/hip_runtime.h"
#include "ntt.h"
#include "ntt_math.h"

namespace kz {

// Two tile sizes, one radix-4 group per thread per stage pair:
//   2^11 elements (72.3 KiB of LDS), 512 threads: two workgroups per CU
//   2^10 elements (36.1 KiB), 256 threads: four workgroups per CU, so a workgroup
//        waiting at a barrier leaves three to fill the SIMDs
// A pass of 2^11 points needs the large tile; shorter passes take the small one.
template <int TileLog>
struct NttTile {
    static constexpr int kLog = TileLog;
    static constexpr int kElems = 1 << TileLog;
    static constexpr int kThreads = kElems / 4;
};

struct NttPass {
    uint32_t logL, T, Q0;
    uint32_t in_b0, in_b1, in_s;
    uint32_t out_b0, out_b1, out_s;
    uint32_t four_step;  // multiply output k of column q0 by ftw[k * Q0 + q0]
    uint32_t scale;      // last pass of an inverse: multiply by n^-1 (canonical)
    uint32_t check;      // flag inputs >= r
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

// LDS word offset of tile element idx: nine words per element plus one pad
// word per 32 elements.  Without the pad, elements 32 apart start on the same
// bank (9 * 32 = 0 mod 32), which the bit-reversed load (consecutive j land
// 32 * k apart), the strided tiles (sub-DFTs L apart) and the first radix-4
// stage pair (stride 4: 4-way) all hit.
__device__ __forceinline__ uint32_t lds_off(uint32_t idx) { return idx * 9u + (idx >> 5); }

__device__ __forceinline__ uint32_t bitrev(uint32_t j, uint32_t bits) {
    return bits ? (__brev(j) >> (32 - bits)) : 0u;
}

// Either tile keeps four waves per SIMD: at most 128 VGPRs.
template <class F, int TileLog>
__global__ void __launch_bounds__(NttTile<TileLog>::kThreads) __attribute__((amdgpu_waves_per_eu(4)))
k_ntt_pass(const uint32_t* __restrict__ in, uint32_t* __restrict__ out, const uint32_t* __restrict__ stw,
           const uint32_t* __restrict__ ftw, const uint32_t* __restrict__ ninv, NttPass d, uint32_t n_mask,
           int* __restrict__ err) {
    static_assert(F::N == 9, "Fr limbs");
    constexpr uint32_t kNttThreads = NttTile<TileLog>::kThreads;
    __shared__ uint32_t sh[NttTile<TileLog>::kElems * F::N + NttTile<TileLog>::kElems / 32];
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
        fe_store_words9<F>(sh + lds_off(b * L + bitrev(j, d.logL)), x);
    }
    if (bad) atomicOr(err, 1);
    __syncthreads();

    // Butterflies: radix-2 DIT stages s = 0 .. logL-1 (stage s pairs elements
    // h = 2^s apart with twiddle w_(2h)^i = w_L^(i << (logL-1-s))), fused two
    // at a time (ntt_radix4), so LDS is read and written once per two stages.
    // An odd logL starts with the twiddle-free stage 0 alone.  Groups never
    // straddle sub-DFTs (4h | L).
    uint32_t s = 0;
    if (d.logL & 1) {
        for (uint32_t p = tid; p < tile / 2; p += kNttThreads) {
            Fe<F> u = fe_load_words9<F>(sh + lds_off(2 * p));
            Fe<F> v = fe_load_words9<F>(sh + lds_off(2 * p + 1));
            ntt_bfly0<F>(u, v);
            fe_store_words9<F>(sh + lds_off(2 * p), u);
            fe_store_words9<F>(sh + lds_off(2 * p + 1), v);
        }
        __syncthreads();
        s = 1;
    }
    for (; s < d.logL; s += 2) {
        const uint32_t h = 1u << s, sh1 = d.logL - 1 - s, sh2 = d.logL - 2 - s;
        for (uint32_t q = tid; q < tile / 4; q += kNttThreads) {
            const uint32_t i = q & (h - 1), e = ((q >> s) << (s + 2)) + i;
            Fe<F> x0 = fe_load_words9<F>(sh + lds_off(e));
            Fe<F> x1 = fe_load_words9<F>(sh + lds_off(e + h));
            Fe<F> x2 = fe_load_words9<F>(sh + lds_off(e + 2 * h));
            Fe<F> x3 = fe_load_words9<F>(sh + lds_off(e + 3 * h));
            const Fe<F> w1 = fe_load_words9<F>(stw + (size_t)(i << sh2) * F::N);
            const Fe<F> w2 = fe_load_words9<F>(stw + (size_t)((i + h) << sh2) * F::N);
            if (s == 0) {
                ntt_radix4<F, true>(x0, x1, x2, x3, w1, w1, w2);
            } else {
                const Fe<F> ws = fe_load_words9<F>(stw + (size_t)(i << sh1) * F::N);
                ntt_radix4<F, false>(x0, x1, x2, x3, ws, w1, w2);
            }
            fe_store_words9<F>(sh + lds_off(e), x0);
            fe_store_words9<F>(sh + lds_off(e + h), x1);
            fe_store_words9<F>(sh + lds_off(e + 2 * h), x2);
            fe_store_words9<F>(sh + lds_off(e + 3 * h), x3);
        }
        __syncthreads();
    }

    const Fe<F> scale = d.scale ? fe_load_words9<F>(ninv) : fe_zero<F>();
    for (uint32_t e = tid; e < tile; e += kNttThreads) {
        uint32_t b, k;
        if (d.out_s == 1) { b = e >> d.logL; k = e & (L - 1); }
        else { b = e % d.T; k = e / d.T; }
        const uint32_t q = q_start + b, q0 = q % d.Q0, q1 = q / d.Q0;
        Fe<F> x = fe_load_words9<F>(sh + lds_off(b * L + k));
        if (d.four_step) x = fe_mul_lazy<F>(x, fe_load_words9<F>(ftw + ((size_t)k * d.Q0 + q0) * F::N));
        else if (d.scale) x = fe_mul<F>(x, scale);
        else x = fe_reduce_wide<F>(x);
        uint32_t w[8];
        fe_to_words<F>(x, w, 8);
        const size_t idx = (size_t)q0 * d.out_b0 + (size_t)q1 * d.out_b1 + (size_t)k * d.out_s;
        uint4* dst = reinterpret_cast<uint4*>(out + idx * 8);
        dst[0] = make_uint4(w[0], w[1], w[2], w[3]);
        dst[1] = make_uint4(w[4], w[5], w[6], w[7]);
    }
}

// ------------------------------------------------------------ twiddles
// w = g^exp (exp = (r-1) >> logn, 8 LE words), inverted for the inverse
// transform; out[b] = w^(2^b) for b < logn, out[logn] = n^-1.  Montgomery form.
template <class F>
__global__ void k_ntt_roots(const uint32_t* __restrict__ exp_words, uint32_t g, uint32_t logn, uint32_t inverse,
                            uint32_t* __restrict__ out) {
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
    for (uint32_t b = 0; b < logn; ++b) {
        fe_store_words9<F>(out + b * F::N, acc);
        acc = fe_mul<F>(acc, acc);
    }
    Fe<F> nn = fe_zero<F>();
    nn.v[logn / kLimbBits] = 1u << (logn % kLimbBits);
    fe_store_words9<F>(out + logn * F::N, fe_inv<F>(fe_to_mont<F>(nn)));
}

// tab[idx] = w^e from the powers pw[b] = w^(2^b): e = idx * mult (Q0 = 0), or
// e = k * q0 * mult for idx = k * Q0 + q0; e taken mod n.
template <class F>
__global__ void __launch_bounds__(256) k_ntt_pow_table(const uint32_t* __restrict__ pw, uint32_t logn, size_t count,
                                                       uint32_t Q0, uint32_t mult, uint32_t* __restrict__ tab) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= count) return;
    const uint64_t n_mask = ((uint64_t)1 << logn) - 1;
    const uint64_t e = Q0 ? ((uint64_t)(idx / Q0) * (idx % Q0) * mult) & n_mask : ((uint64_t)idx * mult) & n_mask;
    Fe<F> acc = fe_one<F>();
    for (uint32_t b = 0; b < logn; ++b)
        if ((e >> b) & 1u) acc = fe_mul<F>(acc, fe_load_words9<F>(pw + b * F::N));
    fe_store_words9<F>(tab + idx * F::N, acc);
}

namespace {

// (r - 1) of each Fr as 4 little-endian 64-bit limbs.
const uint64_t kRm1[2][4] = {
    {0xFFFFFFFF00000000ull, 0x53BDA402FFFE5BFEull, 0x3339D80809A1D805ull, 0x73EDA753299D7D48ull},  // BLS12-381
    {0x43E1F593F0000000ull, 0x2833E84879B97091ull, 0xB85045B68181585Dull, 0x30644E72E131A029ull},  // BN254
};
const uint32_t kGen[2] = {7, 5};
const int kTwoAdicity[2] = {32, 28};

// Every table is built on first use and kept (captured graphs hold their
// addresses), except four-step tables the tuner built for splits it did not
// keep (ntt_drop_tables_since).
struct NttTable {
    DevBuf buf;
    size_t seq = 0;
};
std::mutex g_ntt_mu;
size_t g_ntt_seq = 0;
std::map<std::tuple<int, int, int>, std::unique_ptr<NttTable>> g_roots;  // (field, logn, inverse)
std::map<std::tuple<int, int, int>, std::unique_ptr<NttTable>> g_stw;    // (field, logL, inverse)
// (field, logn, inverse, logL, Q0, mult)
std::map<std::tuple<int, int, int, int, uint32_t, uint32_t>, std::unique_ptr<NttTable>> g_ftw;

template <class F>
void launch_pow_table(const uint32_t* pw, int logn, size_t count, uint32_t Q0, uint32_t mult, uint32_t* tab,
                      hipStream_t s) {
    hipLaunchKernelGGL(k_ntt_pow_table<F>, dim3((unsigned)ceil_div(count, 256)), dim3(256), 0, s, pw,
                       (uint32_t)logn, count, Q0, mult, tab);
}

// under g_ntt_mu
const uint32_t* roots_locked(int field, int logn, bool inverse, hipStream_t s) {
    auto key = std::make_tuple(field, logn, inverse ? 1 : 0);
    auto it = g_roots.find(key);
    if (it != g_roots.end()) return it->second->buf.as<uint32_t>();
    auto t = std::make_unique<NttTable>();
    t->seq = g_ntt_seq++;
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
    DevBuf expw;
    expw.ensure(sizeof ew);
    t->buf.ensure((size_t)(logn + 1) * 9 * 4);
    KZ_HIP(hipMemcpyAsync(expw.ptr, ew, sizeof ew, hipMemcpyHostToDevice, s));
    if (field == 0)
        hipLaunchKernelGGL(k_ntt_roots<BlsFr>, dim3(1), dim3(64), 0, s, expw.as<uint32_t>(), kGen[0], (uint32_t)logn,
                           (uint32_t)inverse, t->buf.as<uint32_t>());
    else
        hipLaunchKernelGGL(k_ntt_roots<BnFr>, dim3(1), dim3(64), 0, s, expw.as<uint32_t>(), kGen[1], (uint32_t)logn,
                           (uint32_t)inverse, t->buf.as<uint32_t>());
    KZ_HIP(hipGetLastError());
    KZ_HIP(hipStreamSynchronize(s));  // the staging copy is from the stack; expw is freed on return
    const uint32_t* p = t->buf.as<uint32_t>();
    g_roots[key] = std::move(t);
    return p;
}

// under g_ntt_mu: w_L^e for e < max(1, L/2)
const uint32_t* stw_locked(int field, int logL, bool inverse, hipStream_t s) {
    auto key = std::make_tuple(field, logL, inverse ? 1 : 0);
    auto it = g_stw.find(key);
    if (it != g_stw.end()) return it->second->buf.as<uint32_t>();
    const uint32_t* pw = roots_locked(field, logL, inverse, s);
    auto t = std::make_unique<NttTable>();
    t->seq = g_ntt_seq++;
    const size_t count = logL ? (size_t)1 << (logL - 1) : 1;
    t->buf.ensure(count * 9 * 4);
    if (field == 0) launch_pow_table<BlsFr>(pw, logL, count, 0, 1, t->buf.as<uint32_t>(), s);
    else launch_pow_table<BnFr>(pw, logL, count, 0, 1, t->buf.as<uint32_t>(), s);
    KZ_HIP(hipGetLastError());
    const uint32_t* p = t->buf.as<uint32_t>();
    g_stw[key] = std::move(t);
    return p;
}

// under g_ntt_mu: entry k * Q0 + q0 = w_n^(mult q0 k) for k < L, q0 < Q0
const uint32_t* ftw_locked(int field, int logn, bool inverse, int logL, uint32_t Q0, uint32_t mult, hipStream_t s) {
    auto key = std::make_tuple(field, logn, inverse ? 1 : 0, logL, Q0, mult);
    auto it = g_ftw.find(key);
    if (it != g_ftw.end()) return it->second->buf.as<uint32_t>();
    const uint32_t* pw = roots_locked(field, logn, inverse, s);
    auto t = std::make_unique<NttTable>();
    t->seq = g_ntt_seq++;
    const size_t count = ((size_t)1 << logL) * Q0;
    t->buf.ensure(count * 9 * 4);
    if (field == 0) launch_pow_table<BlsFr>(pw, logn, count, Q0, mult, t->buf.as<uint32_t>(), s);
    else launch_pow_table<BnFr>(pw, logn, count, Q0, mult, t->buf.as<uint32_t>(), s);
    KZ_HIP(hipGetLastError());
    const uint32_t* p = t->buf.as<uint32_t>();
    g_ftw[key] = std::move(t);
    return p;
}

// Tuned pass splits per (field, log2 n): log2 of the sub-transform lengths of
// passes 1..3 (unused passes 0); all zero = the built-in split.
std::mutex g_sched_mu;
int g_sched[2][kNttMaxLog + 1][3];

// One pass up to 2^11; beyond, balanced passes of at most 2^10 (the small
// tile): four 4-wave workgroups per CU hide barrier and LDS waits better than
// the large tile's two 8-wave ones, which measured slower than a third pass.
void builtin_split(int logn, int (&p)[3]) {
    p[0] = p[1] = p[2] = 0;
    if (logn <= kNttPassMaxLog) p[0] = logn;
    else if (logn <= 20) { p[0] = (logn + 1) / 2; p[1] = logn - p[0]; }
    else { p[0] = (logn + 2) / 3; p[1] = (logn - p[0] + 1) / 2; p[2] = logn - p[0] - p[1]; }
}

int tile_log_of(int logL) { return logL > 10 ? 11 : 10; }

struct NttPlan {
    NttPass pass[3];
    const uint32_t* stw[3];
    const uint32_t* ftw[3];
    const uint32_t* ninv;
    int np;
};

// Pass descriptors and their tables (built here on first use: allocates and
// may synchronise the stream).
void plan(int field, int logn, bool inverse, hipStream_t s, NttPlan& P) {
    int sp[3];
    ntt_schedule_of(field, logn, sp[0], sp[1], sp[2]);
    // la, lb, lc: lengths of the column pass, the inner column pass and the
    // row pass (la = lb = 0 for one pass, lb = 0 for two)
    int la, lb, lc, np;
    if (sp[2]) { la = sp[0]; lb = sp[1]; lc = sp[2]; np = 3; }
    else if (sp[1]) { la = sp[0]; lb = 0; lc = sp[1]; np = 2; }
    else { la = 0; lb = 0; lc = sp[0]; np = 1; }
    const uint32_t A = 1u << la, B = 1u << lb, C = 1u << lc, M = B * C;
    // sub-DFTs per workgroup: fill the pass's tile, but never more than the
    // n / L sub-DFTs the pass has
    auto T = [logn](int logL) {
        const uint32_t fit = (uint32_t)(1u << tile_log_of(logL)) >> logL, have = 1u << (logn - logL);
        return fit < have ? fit : have;
    };
    std::lock_guard<std::mutex> g(g_ntt_mu);
    const size_t seq0 = g_ntt_seq;
    int k = 0;
    if (np >= 2) {
        NttPass p{};
        p.logL = la; p.T = T(la); p.Q0 = M;
        p.in_b0 = 1; p.in_b1 = 0; p.in_s = M;
        p.out_b0 = 1; p.out_b1 = 0; p.out_s = M;
        p.four_step = 1;
        P.stw[k] = stw_locked(field, la, inverse, s);
        P.ftw[k] = ftw_locked(field, logn, inverse, la, M, 1, s);
        P.pass[k++] = p;
    }
    if (np == 3) {
        NttPass p{};
        p.logL = lb; p.T = T(lb); p.Q0 = C;
        p.in_b0 = 1; p.in_b1 = M; p.in_s = C;
        p.out_b0 = 1; p.out_b1 = M; p.out_s = C;
        p.four_step = 1;
        P.stw[k] = stw_locked(field, lb, inverse, s);
        P.ftw[k] = ftw_locked(field, logn, inverse, lb, C, A, s);
        P.pass[k++] = p;
    }
    NttPass p{};
    p.logL = lc; p.T = T(lc); p.Q0 = A;
    p.in_b0 = M; p.in_b1 = C; p.in_s = 1;
    p.out_b0 = 1; p.out_b1 = A; p.out_s = A * B;
    p.scale = inverse ? 1 : 0;
    P.stw[k] = stw_locked(field, lc, inverse, s);
    P.ftw[k] = nullptr;
    P.pass[k++] = p;
    P.pass[0].check = 1;
    P.np = np;
    P.ninv = roots_locked(field, logn, inverse, s) + (size_t)logn * 9;
    // a table built here is visible to every stream once its kernel has run
    if (g_ntt_seq != seq0) KZ_HIP(hipStreamSynchronize(s));
}

}  // namespace

size_t ntt_table_mark() {
    std::lock_guard<std::mutex> g(g_ntt_mu);
    return g_ntt_seq;
}

void ntt_drop_tables_since(size_t mark) {
    std::lock_guard<std::mutex> g(g_ntt_mu);
    for (auto it = g_ftw.begin(); it != g_ftw.end();) {
        if (it->second->seq >= mark) it = g_ftw.erase(it);
        else ++it;
    }
}

bool ntt_split_valid(int logn, int p1, int p2, int p3) {
    if (logn < 0 || logn > kNttMaxLog) return false;
    if (p1 < 0 || p2 < 0 || p3 < 0 || p1 > kNttPassMaxLog || p2 > kNttPassMaxLog || p3 > kNttPassMaxLog)
        return false;
    if (p1 + p2 + p3 != logn) return false;
    if (p3 && (!p1 || !p2)) return false;  // zeros only trailing
    if (p2 && !p1) return false;
    return true;
}

void ntt_set_schedule(int field, int logn, int p1, int p2, int p3) {
    if (field < 0 || field > 1 || logn < 0 || logn > kNttMaxLog) throw std::invalid_argument("NTT size class");
    if (!(p1 == 0 && p2 == 0 && p3 == 0) && !ntt_split_valid(logn, p1, p2, p3))
        throw std::invalid_argument("NTT pass split must sum to log2 n with passes of at most 2^11");
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
    NttPlan P;
    plan(field, logn, inverse, s, P);
}

void ntt_device_batch(int field, int logn, bool inverse, const uint32_t* d_in, uint32_t* d_out, uint32_t* d_tmp,
                      size_t count, int* d_err, hipStream_t s) {
    if (field < 0 || field > 1 || logn < 0 || logn > kNttMaxLog || logn > kTwoAdicity[field])
        throw std::invalid_argument("NTT size out of range");
    if (count == 0) return;
    if (count > 65535) throw std::invalid_argument("NTT batch larger than 65535 transforms");
    NttPlan P;
    plan(field, logn, inverse, s, P);
    const uint32_t n = 1u << logn;
    for (int i = 0; i < P.np; ++i) {
        const uint32_t* src = (i == 0) ? d_in : d_tmp;
        uint32_t* dst = (i == P.np - 1) ? d_out : d_tmp;
        const NttPass& p = P.pass[i];
        const unsigned grid = (n >> p.logL) / p.T;  // exact: powers of two, T <= n / L
        const dim3 g(grid, (unsigned)count);
        if (tile_log_of((int)p.logL) == 11) {
            const dim3 b(NttTile<11>::kThreads);
            if (field == 0)
                hipLaunchKernelGGL((k_ntt_pass<BlsFr, 11>), g, b, 0, s, src, dst, P.stw[i], P.ftw[i], P.ninv, p,
                                   n - 1, d_err);
            else
                hipLaunchKernelGGL((k_ntt_pass<BnFr, 11>), g, b, 0, s, src, dst, P.stw[i], P.ftw[i], P.ninv, p,
                                   n - 1, d_err);
        } else {
            const dim3 b(NttTile<10>::kThreads);
            if (field == 0)
                hipLaunchKernelGGL((k_ntt_pass<BlsFr, 10>), g, b, 0, s, src, dst, P.stw[i], P.ftw[i], P.ninv, p,
                                   n - 1, d_err);
            else
                hipLaunchKernelGGL((k_ntt_pass<BnFr, 10>), g, b, 0, s, src, dst, P.stw[i], P.ftw[i], P.ninv, p,
                                   n - 1, d_err);
        }
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
