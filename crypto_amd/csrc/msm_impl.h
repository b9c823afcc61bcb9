// Curve-templated kernels of the Pippenger MSM.
//
// Same mathematics as the reference's CPU body
// (gpukit/cpp/cpu_reference/multi_pippenger.cpp:104-181): signed c-bit digits,
// 2^(c-1) buckets per window, bucket sums weighted by (bucket + 1), windows
// combined MSW -> LSW.  The schedule is the GPU's own:
//
//   k_digits (count)   atomic histogram over global bucket ids (msm_digits.hip)
//   scan               bucket offsets (scan.hip)
//   k_digits (scatter) point indices into bucket order
//   k_accumulate       one thread per fixed-length chunk of the sorted list:
//                      mixed XYZZ additions; the chunk holding a bucket's first
//                      entry writes the bucket sum, a chunk whose first entries
//                      continue an earlier bucket writes them as its carry
//   k_reduce_seg       per 16-bucket segment: folds the carries into the
//                      bucket sums, then (sum, weighted sum) by running sums
//   k_reduce_level     log-depth levels (LDS suffix scan + tree) down to one
//                      sum per window
//   (window combine: k_combine_row, msm_combine_row.hip)
//
// The accumulation inlines its field products (throughput-bound, one compact
// loop).  The reduction kernels use quad-cooperative point operations
// (curve_quad.h), the combine row-distributed ones (field_row.h); the
// reduction calls one out-of-line field
// multiply (MulCall) so its code stays inside the instruction cache, and keeps
// few point-operation sites with `#pragma unroll 1` loops.
#pragma once

#include "msm_internal.h"
#include "point_ops.h"

namespace kz {

// ------------------------------------------------------------ wire -> device
template <class C>
__global__ void __launch_bounds__(128)
k_points_from_wire(const uint8_t* __restrict__ wire, uint32_t* __restrict__ pts, uint8_t* __restrict__ inf,
                   uint32_t n) {
    using F = typename C::Fp;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    constexpr int WORDS = C::FB / 4;  // words per coordinate on the wire
    const uint32_t* w = reinterpret_cast<const uint32_t*>(wire + (size_t)i * 2 * C::FB);
    uint32_t xw[WORDS], yw[WORDS];
    uint32_t any = 0;
#pragma unroll
    for (int k = 0; k < WORDS; ++k) {
        const uint32_t wx = w[WORDS - 1 - k];
        const uint32_t wy = w[2 * WORDS - 1 - k];
        xw[k] = __builtin_bswap32(wx);
        yw[k] = __builtin_bswap32(wy);
        any |= wx | wy;
    }
    Fe<F> x = fe_from_words<F>(xw, WORDS), y = fe_from_words<F>(yw, WORDS);
    // The reference reads coordinates mod p (to_mont_fp reduces any value
    // below 2^(32N)); reduce non-canonical coordinates the same way.
    while (!fe_is_canonical<F>(x)) fe_reduce_once<F>(x);
    while (!fe_is_canonical<F>(y)) fe_reduce_once<F>(y);
    Affine<F> a;
    a.x = fe_to_mont<F>(x);
    a.y = fe_to_mont<F>(y);
    affine_store<F>(pts, i, a);
    inf[i] = any ? 0 : 1;
}

// ------------------------------------------------------------ accumulation
template <class F>
__device__ __forceinline__ typename PointOps<F>::Aff load_signed_point(const uint32_t* __restrict__ pts,
                                                                      uint32_t entry) {
    using P = PointOps<F>;
    const typename P::Aff q = P::load_aff(pts, entry & 0x7FFFFFFFu);
    return (entry >> 31) ? P::neg_aff(q) : q;
}

// Largest g with offsets[g] <= pos < offsets[g + 1] (pos < offsets[nb]).
__device__ __forceinline__ uint32_t bucket_of(const uint32_t* __restrict__ offsets, uint32_t nb, uint32_t pos) {
    uint32_t lo = 0, hi = nb;  // invariant: offsets[lo] <= pos < offsets[hi]
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (offsets[mid] <= pos) lo = mid;
        else hi = mid;
    }
    return lo;
}

template <class F>
__global__ void __launch_bounds__(64, 3)
k_accumulate(const uint32_t* __restrict__ entries, const uint32_t* __restrict__ offsets, uint32_t nb,
             uint32_t chunk, const uint32_t* __restrict__ pts, uint32_t* __restrict__ bsum,
             uint32_t* __restrict__ carry) {
    using P = PointOps<F>;
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t total = offsets[nb];
    const uint32_t start = j * chunk;
    if (start >= total) return;
    const uint32_t end = min(start + chunk, total);
    uint32_t g = bucket_of(offsets, nb, start);
    uint32_t next = offsets[g + 1];
    bool is_carry = offsets[g] < start;
    typename P::Pt acc = P::from_aff(load_signed_point<F>(pts, entries[start]));
    for (uint32_t e = start + 1; e < end; ++e) {
        const typename P::Aff q = load_signed_point<F>(pts, entries[e]);
        if (e == next) {
            if (is_carry) P::store(carry, j, P::canon(acc));
            else P::store(bsum, g, P::canon(acc));
            do {
                ++g;
                next = offsets[g + 1];
            } while (next == e);
            is_carry = false;
            acc = P::from_aff(q);
        } else {
            P::madd_acc(acc, q);
        }
    }
    if (is_carry) P::store(carry, j, P::canon(acc));
    else P::store(bsum, g, P::canon(acc));
}

// ------------------------------------------------------------ carry fold
// The carries of a bucket whose entries span chunks j0..j1 (j0 = the chunk
// holding its first entry) sit in carry[j0+1 .. j1]; they must all end up in
// carry[j0+1], the head, which the reduction adds to the bucket sum.  With
// random scalars most buckets have no carry or one; runs are long when a
// bucket is large -- the top window's few buckets, GLV's carry window, or the
// single shared bucket set of a window table (about rows * n / 2^(c-1)
// entries per bucket).  So:
//   k_carry_runs   one quad per bucket reads its run [head, head + len) off
//                  the offsets; runs of at most kShortRun carries are folded
//                  serially by that quad (quad-cooperative additions); a
//                  longer run is recorded as {head, len} (and the longest
//                  length is kept);
//   k_list_runs    a workgroup per recorded run reserves len slots of a
//                  compact list (one atomic) and its threads write {chunk j,
//                  position r in the run, len} for the run's chunks (one run
//                  can hold most of a window's chunks);
//   k_fold_groups  tree of fan-in 64 over the listed chunks: at level L (stride
//                  S = 64^L) the entry at run position r = 0 mod 64 S leads a
//                  group, whose workgroup (one quad per element) sums the
//                  carries at positions r + e S, e < 64, through a 6-step LDS
//                  tree into carry[j].  Levels past the longest run return at
//                  once; a 2^21-point MSM needs 3 launches instead of 17
//                  pairwise ones.  Grid-stride over the list.
// ctl[0] = list length, ctl[1] = longest listed run, ctl[2] = recorded runs
// (all zeroed before).
// (Templated only so that each per-curve translation unit owns its copy.)
constexpr uint32_t kShortRun = kFoldShortRun;

template <class F>
__global__ void __launch_bounds__(256)
k_carry_runs(const uint32_t* __restrict__ offsets, uint32_t nb, uint32_t chunk, uint2* __restrict__ runs,
             uint32_t* __restrict__ ctl, uint32_t* __restrict__ carry) {
    using P = PointOps<F>;
    const uint32_t g = (blockIdx.x * blockDim.x + threadIdx.x) >> 2;
    if (g >= nb) return;  // whole quads leave together
    const uint32_t lo = offsets[g], hi = offsets[g + 1];
    if (hi <= lo) return;  // empty bucket
    const uint32_t head = lo / chunk + 1, len = (hi - 1) / chunk + 1 - head;
    if (len <= 1) return;  // no carry, or one already in place
    if (len <= kShortRun) {
        typename P::Pt acc = P::load(carry, head);
#pragma unroll 1
        for (uint32_t k = 1; k < len; ++k) acc = P::template add_q<MulCallLazy>(acc, P::load(carry, head + k));
        if (quad_lane() == 0) P::store(carry, head, P::canon(acc));
        return;
    }
    if (quad_lane() == 0) {
        runs[atomicAdd(&ctl[2], 1u)] = make_uint2(head, len);
        atomicMax(&ctl[1], len);
    }
}

template <class F>
__global__ void __launch_bounds__(256)
k_list_runs(const uint2* __restrict__ runs, uint4* __restrict__ list, uint32_t* __restrict__ ctl) {
    __shared__ uint32_t base;
    const uint32_t nruns = ctl[2];
#pragma unroll 1
    for (uint32_t r = blockIdx.x; r < nruns; r += gridDim.x) {
        const uint2 run = runs[r];  // {head, len}
        __syncthreads();            // the previous run's readers of `base` are done
        if (threadIdx.x == 0) base = atomicAdd(&ctl[0], run.y);
        __syncthreads();
        for (uint32_t k = threadIdx.x; k < run.y; k += blockDim.x) list[base + k] = make_uint4(run.x + k, k, run.y, 0u);
    }
}

constexpr uint32_t kFoldFan = 64;

template <class F>
__global__ void __launch_bounds__(4 * kFoldFan)
k_fold_groups(const uint4* __restrict__ list, const uint32_t* __restrict__ ctl, uint32_t stride,
              uint32_t* __restrict__ carry) {
    using P = PointOps<F>;
    __shared__ typename P::Pt sh[kFoldFan];
    if (stride >= ctl[1]) return;  // every run is one element at this level
    const uint32_t count = ctl[0], q = threadIdx.x >> 2;
    const uint64_t span = (uint64_t)stride * kFoldFan;
#pragma unroll 1
    for (uint32_t i = blockIdx.x; i < count; i += gridDim.x) {  // one entry per workgroup step
        const uint4 en = list[i];                                 // {chunk j, run position r, run length}
        if (en.y % span != 0) continue;                           // not a group leader at this level
        const uint64_t pos = en.y + (uint64_t)q * stride;
        typename P::Pt p = pos < en.z ? P::load(carry, en.x + (size_t)q * stride) : P::inf();
#pragma unroll 1
        for (uint32_t st = 1; st < kFoldFan; st <<= 1) {
            __syncthreads();  // the previous step's readers of sh are done
            if ((q & (2 * st - 1)) == st && quad_lane() == 0) sh[q] = p;
            __syncthreads();
            if ((q & (2 * st - 1)) == 0) p = P::template add_q<MulCallLazy>(p, sh[q + st]);
        }
        if (q == 0 && quad_lane() == 0) P::store(carry, en.x, P::canon(p));
    }
}

// ---------------------------------------------------------- bucket reduce
// Window sum W = sum_b (b+1) S_b over the window's bucket sums, in levels.
// (The chains of point additions below run on lazily reduced coordinates,
// MulCallLazy, below 8p / 4p / 2p, and store canonical ones.)
//
//   k_reduce_seg    per segment t of `seg` buckets (a power of two dividing
//                   the bucket count; one quad each): adds each bucket's folded
//                   carry run to its sum, then by running sums from the top
//                   A_t = sum_l S_{seg t+l},  B_t = sum_l (l+1) S_{seg t+l},
//                   so  W = sum_t [B_t + seg t A_t].
//   k_reduce_level  generic step.  A window's inputs (A_i, B_i), i < n, with
//                   W = sum_i [B_i + m i A_i]  (m = 2^lm)  are cut into parts of
//                   kLevelQuads = 128 consecutive inputs; part g yields
//                   A'_g = sum_k A_{128g+k},  B'_g = sum_k [B_{128g+k} + m k A_{128g+k}],
//                   and W = sum_g [B'_g + 128 m g A'_g]: the same form with 128x
//                   fewer inputs.  Levels repeat until one part is left, whose
//                   B' is W.
// Inside a part (one workgroup, one quad per input): sum_k k A_k =
// sum_{k>=1} R_k with the suffix sums R_k = sum_{j>=k} A_j (LDS scan), so
// B' = sum_k term_k with term_0 = B_0, term_k = B_k + m R_k (LDS tree), and
// A' = R_0.  The running sums cost 3 point additions per bucket; the levels,
// which see 16x and more fewer inputs, trade extra additions for depth
// log2(128) per scan/tree instead of a serial pass over a window.
constexpr int kLevelQuads = (int)kReduceFanIn;
constexpr int kLevelThreads = 4 * kLevelQuads;

// Two waves per SIMD: unbounded, the BLS12-381 form needs ~315 registers
// (one wave per SIMD) and a 2^20 MSM's ~2300 segment waves run in 2-3
// rounds; capped at 256 it spills ~60 registers and runs in one round
// (reduce stage 1.10 -> 1.03 ms for a 2^20 GLV MSM, 1.53 -> 1.39 ms plain).
template <class F>
__global__ void __launch_bounds__(256, 2)
k_reduce_seg(const uint32_t* __restrict__ offsets, uint32_t chunk, const uint32_t* __restrict__ bsum,
             const uint32_t* __restrict__ carry, uint32_t nseg, uint32_t seg, uint32_t* __restrict__ segA,
             uint32_t* __restrict__ segB) {
    using P = PointOps<F>;
    const uint32_t t = (blockIdx.x * blockDim.x + threadIdx.x) >> 2;
    if (t >= nseg) return;  // whole quads leave together
    typename P::Pt run = P::inf(), tot = P::inf();
#pragma unroll 1
    for (int l = (int)seg - 1; l >= 0; --l) {
        const uint32_t gb = t * seg + (uint32_t)l;
        const uint32_t off = offsets[gb], end = offsets[gb + 1];
        // the bucket sum and its head carry go into run one after the other,
        // so at most three points are live (run, tot and one loaded point)
        if (end > off) {
            run = P::template add_q<MulCallLazy>(run, P::load(bsum, gb));
            const uint32_t head = off / chunk + 1;
            if ((end - 1) / chunk >= head) run = P::template add_q<MulCallLazy>(run, P::load(carry, head));
        }
        tot = P::template add_q<MulCallLazy>(tot, run);
    }
    if (quad_lane() == 0) {
        P::store(segA, t, P::canon(run));
        P::store(segB, t, P::canon(tot));
    }
}

// tab = null: every window has uni_n inputs at w * uni_n and uni_parts parts,
// written at w * uni_parts + g.  Else per window (in_base, n_in, out_base,
// wg_base), workgroup b serving the last window with wg_base <= b.
template <class F>
__global__ void __launch_bounds__(kLevelThreads)
k_reduce_level(const uint32_t* __restrict__ inA, const uint32_t* __restrict__ inB,
               const RedLevel* __restrict__ tab, uint32_t ntab, uint32_t uni_n, uint32_t uni_parts, uint32_t lm,
               uint32_t* __restrict__ outA, uint32_t* __restrict__ outB) {
    using P = PointOps<F>;
    __shared__ typename P::Pt sh[kLevelQuads];
    uint32_t in_base, n_in, out_idx, g;
    if (tab) {
        uint32_t lo = 0, hi = ntab;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (tab[mid].wg_base <= blockIdx.x) lo = mid;
            else hi = mid;
        }
        const RedLevel d = tab[lo];
        g = blockIdx.x - d.wg_base;
        in_base = d.in_base;
        n_in = d.n_in;
        out_idx = d.out_base + g;
    } else {
        const uint32_t w = blockIdx.x / uni_parts;
        g = blockIdx.x - w * uni_parts;
        in_base = w * uni_n;
        n_in = uni_n;
        out_idx = blockIdx.x;
    }
    const uint32_t k = threadIdx.x >> 2;
    const uint32_t first = g * kLevelQuads;
    const uint32_t cnt = min(n_in - first, (uint32_t)kLevelQuads);
    uint32_t span = 1;
    while (span < cnt) span <<= 1;
    const bool lead = quad_lane() == 0;

    typename P::Pt R = P::inf(), term = P::inf();
    if (k < cnt) {
        R = P::load(inA, (size_t)in_base + first + k);
        term = P::load(inB, (size_t)in_base + first + k);
    }
    // suffix scan R_k = sum_{j >= k} A_j over the part
    if (lead) sh[k] = R;
    __syncthreads();
#pragma unroll 1
    for (uint32_t s = 1; s < span; s <<= 1) {
        const typename P::Pt other = (k + s < span) ? sh[k + s] : P::inf();
        __syncthreads();
        if (k < span) R = P::template add_q<MulCallLazy>(R, other);
        if (lead) sh[k] = R;
        __syncthreads();
    }
    // term_k = B_k + m R_k (k > 0)
    if (k > 0 && k < span) {
        typename P::Pt mR = R;
#pragma unroll 1
        for (uint32_t j = 0; j < lm; ++j) mR = P::template dbl_q<MulCallLazy>(mR);
        term = P::template add_q<MulCallLazy>(term, mR);
    }
    if (lead) sh[k] = term;
    __syncthreads();
#pragma unroll 1
    for (uint32_t s = span >> 1; s > 0; s >>= 1) {
        if (k < s) {
            const typename P::Pt v = P::template add_q<MulCallLazy>(sh[k], sh[k + s]);
            if (lead) sh[k] = v;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        P::store(outA, out_idx, P::canon(R));
        P::store(outB, out_idx, P::canon(sh[0]));
    }
}

// --------------------------------------------------------- combine + output
template <class C>
__device__ void write_compressed(const Fe<typename C::Fp>& x_canon, const Fe<typename C::Fp>& y_canon,
                                 uint8_t* __restrict__ comp) {
    using F = typename C::Fp;
    if constexpr (C::ID == 0) {  // secp256k1: no compressed form in the record
        for (int k = 0; k < C::FB; ++k) comp[k] = 0;
        return;
    }
    const bool larger = fe_gt<F>(y_canon, fe_neg<F>(y_canon));  // y > p - y
    fe_to_be_bytes<F>(x_canon, comp, C::FB);
    if (C::ID == 2) comp[0] |= 0x80 | (larger ? 0x20 : 0);
    else comp[0] |= larger ? 0xC0 : 0x80;
}

template <class C>
__device__ void write_infinity(uint8_t* __restrict__ out) {
    for (int k = 0; k < 3 * C::FB; ++k) out[k] = 0;
    out[2 * C::FB] = (C::ID == 2) ? 0xC0 : (C::ID == 1) ? 0x40 : 0x00;
}

template <class C>
__device__ void write_output(const Xyzz<typename C::Fp>& r, uint8_t* __restrict__ out) {
    using F = typename C::Fp;
    Affine<F> a;
    bool inf;
    xyzz_to_affine<F>(r, a, inf);
    if (inf) {
        write_infinity<C>(out);
        return;
    }
    const Fe<F> x = fe_from_mont<F>(a.x), y = fe_from_mont<F>(a.y);
    fe_to_be_bytes<F>(x, out, C::FB);
    fe_to_be_bytes<F>(y, out + C::FB, C::FB);
    write_compressed<C>(x, y, out + 2 * C::FB);
}

// Banderwagon record: affine x || y (the identity as (0, 1), as the
// reference's serialize_uncompressed writes it, element.cpp:286-300), then the
// 32-byte compressed form of serialize_compressed (element.cpp:266-284): x,
// negated unless y > (q - 1) / 2.
template <class C>
__device__ void write_output(const EdExt<typename C::Fp>& r, uint8_t* __restrict__ out) {
    using F = typename C::Fp;
    EdAff<F> a;
    bool inf;
    ed_to_affine<F>(r, a, inf);
    Fe<F> x = fe_zero<F>(), y = fe_zero<F>();
    y.v[0] = 1;
    if (!inf) {
        x = fe_from_mont<F>(a.x);
        y = fe_from_mont<F>(a.y);
    }
    fe_to_be_bytes<F>(x, out, C::FB);
    fe_to_be_bytes<F>(y, out + C::FB, C::FB);
    const bool largest = fe_gt<F>(y, fe_neg<F>(y));  // y > q - y, i.e. y > (q - 1) / 2
    fe_to_be_bytes<F>(largest ? x : fe_neg<F>(x), out + 2 * C::FB, C::FB);
}

// ------------------------------------------------------------ window tables
// Rows of a window table: row w holds 2^(c w) P_i at w n + i (row 0 = the
// points), as affine Montgomery records; one thread per point walks the rows
// by c doublings each and normalises every row point on its own.  The curves'
// groups have odd order, so doubling a non-identity point never gives the
// identity.  Identity points (inf[i]) are copied as they are into every row:
// MSMs skip them through the same flags.
template <class C>
__global__ void __launch_bounds__(64)
k_table_rows(const uint32_t* __restrict__ pts, const uint8_t* __restrict__ inf, uint32_t n, uint32_t c,
             uint32_t nwin, uint32_t* __restrict__ table) {
    using F = typename C::Fp;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Affine<F> a = affine_load<F>(pts, i);
    affine_store<F>(table, i, a);
    const bool id = inf && inf[i];
    Xyzz<F> q = xyzz_from_affine<F>(a);
#pragma unroll 1
    for (uint32_t w = 1; w < nwin; ++w) {
        Affine<F> r = a;
        if (!id) {
#pragma unroll 1
            for (uint32_t k = 0; k < c; ++k) q = xyzz_dbl<F>(q);
            bool is_inf;
            xyzz_to_affine<F>(q, r, is_inf);
        }
        affine_store<F>(table, (size_t)w * n + i, r);
    }
}

template <class C>
__global__ void k_compress_point(const uint8_t* __restrict__ wire, uint8_t* __restrict__ out) {
    using F = typename C::Fp;
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    uint32_t any = 0;
    for (int k = 0; k < 2 * C::FB; ++k) any |= wire[k];
    if (!any) {
        for (int k = 0; k < C::FB; ++k) out[k] = 0;
        out[0] = (C::ID == 2) ? 0xC0 : 0x40;
        return;
    }
    const Fe<F> x = fe_from_be_bytes<F>(wire, C::FB);
    const Fe<F> y = fe_from_be_bytes<F>(wire + C::FB, C::FB);
    write_compressed<C>(x, y, out);
}

}  // namespace kz
