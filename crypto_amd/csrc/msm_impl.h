// Curve-templated kernels of the Pippenger MSM and their orchestration.
// Included by exactly one translation unit per curve (msm_bls.hip, msm_bn.hip).
//
// Same mathematics as the reference's CPU body
// (gpukit/cpp/cpu_reference/multi_pippenger.cpp:104-181): signed c-bit digits,
// 2^(c-1) buckets per window, bucket sums weighted by (bucket + 1), windows
// combined MSW -> LSW.  The schedule is the GPU's own:
//
//   k_digits (count)   atomic histogram over global bucket ids (msm_digits.hip)
//   scan               bucket offsets (scan.h)
//   k_digits (scatter) point indices into bucket order
//   k_accumulate       one thread per fixed-length chunk of the sorted list:
//                      mixed XYZZ additions; the chunk holding a bucket's first
//                      entry writes the bucket sum, a chunk whose first entries
//                      continue an earlier bucket writes them as its carry
//   k_fixup            folds carries into buckets spanning several chunks
//   k_reduce_seg       per 16-bucket segment: (sum, weighted sum) by running sums
//   k_reduce_window    one workgroup per window: LDS suffix scan + tree
//   k_combine          Horner over windows, affine normalisation, compression
#pragma once

#include "curve.h"
#include "msm_internal.h"
#include "scan.h"

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
    Fe<F> x = fe_zero<F>(), y = fe_zero<F>();
    uint32_t any = 0;
#pragma unroll
    for (int k = 0; k < WORDS; ++k) {
        const uint32_t wx = w[WORDS - 1 - k];
        const uint32_t wy = w[2 * WORDS - 1 - k];
        x.v[k] = __builtin_bswap32(wx);
        y.v[k] = __builtin_bswap32(wy);
        any |= wx | wy;
    }
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
__device__ __forceinline__ Affine<F> load_signed_point(const uint32_t* __restrict__ pts, uint32_t entry) {
    Affine<F> q = affine_load<F>(pts, entry & 0x7FFFFFFFu);
    if (entry >> 31) q.y = fe_neg<F>(q.y);
    return q;
}

template <class F>
__global__ void __launch_bounds__(64)
k_accumulate(const uint32_t* __restrict__ entries, const uint32_t* __restrict__ ebucket,
             const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ d_total, uint32_t chunk,
             const uint32_t* __restrict__ pts, uint32_t* __restrict__ bsum, uint32_t* __restrict__ carry) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t total = *d_total;
    const uint32_t start = j * chunk;
    if (start >= total) return;
    const uint32_t end = min(start + chunk, total);
    uint32_t g = ebucket[start];
    bool is_carry = offsets[g] < start;
    Xyzz<F> acc = xyzz_from_affine<F>(load_signed_point<F>(pts, entries[start]));
    for (uint32_t e = start + 1; e < end; ++e) {
        const uint32_t ge = ebucket[e];
        const Affine<F> q = load_signed_point<F>(pts, entries[e]);
        if (ge != g) {
            if (is_carry) xyzz_store<F>(carry, j, acc);
            else xyzz_store<F>(bsum, g, acc);
            g = ge;
            is_carry = false;
            acc = xyzz_from_affine<F>(q);
        } else {
            xyzz_madd<F>(acc, q);
        }
    }
    if (is_carry) xyzz_store<F>(carry, j, acc);
    else xyzz_store<F>(bsum, g, acc);
}

template <class F>
__global__ void __launch_bounds__(64)
k_fixup(const uint32_t* __restrict__ counts, const uint32_t* __restrict__ offsets, uint32_t chunk, uint32_t nb,
        uint32_t* __restrict__ bsum, const uint32_t* __restrict__ carry) {
    const uint32_t gb = blockIdx.x * blockDim.x + threadIdx.x;
    if (gb >= nb) return;
    const uint32_t cnt = counts[gb];
    if (cnt == 0) {
        xyzz_store<F>(bsum, gb, xyzz_inf<F>());
        return;
    }
    const uint32_t off = offsets[gb];
    const uint32_t j0 = off / chunk, j1 = (off + cnt - 1) / chunk;
    if (j1 == j0) return;
    Xyzz<F> acc = xyzz_load<F>(bsum, gb);
    for (uint32_t j = j0 + 1; j <= j1; ++j) acc = xyzz_add_ol<F>(acc, xyzz_load<F>(carry, j));
    xyzz_store<F>(bsum, gb, acc);
}

// ---------------------------------------------------------- bucket reduce
constexpr int kSeg = 16;  // buckets per first-level segment

// Segment t (of all windows' segments): A = sum_l S_l, Bw = sum_l (l+1) S_l.
template <class F>
__global__ void __launch_bounds__(64)
k_reduce_seg(const uint32_t* __restrict__ bsum, uint32_t nseg, uint32_t* __restrict__ segA,
             uint32_t* __restrict__ segB) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nseg) return;
    Xyzz<F> run = xyzz_inf<F>(), tot = xyzz_inf<F>();
    // running sums from the top bucket down; one add site per step
    for (int step = 2 * kSeg - 1; step >= 0; --step) {
        const bool load = (step & 1) != 0;
        Xyzz<F> a = load ? run : tot;
        Xyzz<F> b = load ? xyzz_load<F>(bsum, (size_t)t * kSeg + (step >> 1)) : run;
        Xyzz<F> r = xyzz_add_ol<F>(a, b);
        if (load) run = r;
        else tot = r;
    }
    xyzz_store<F>(segA, t, run);
    xyzz_store<F>(segB, t, tot);
}

template <class F>
__device__ __forceinline__ Xyzz<F> dbl_n(Xyzz<F> p, int k) {
    for (int i = 0; i < k; ++i) p = xyzz_dbl_ol<F>(p);
    return p;
}

constexpr int kReduceThreads = 256;

// One workgroup per window over its T1 = B/16 segments:
//   window = sum_t Bw_t + 16 (U - V),  U = sum_t (t+1) A_t,  V = sum_t A_t.
// Thread k owns L2 = T1/256 consecutive segments (one if T1 < 256):
//   U = sum_k u_k + L2 (sum_k R_k - V),  R_k = sum_{j >= k} a_k (suffix scan),
// so window = sum_k [b_k + 16 (u_k + L2 R_k)] - 16 (L2 + 1) V.
template <class F>
__global__ void __launch_bounds__(kReduceThreads)
k_reduce_window(const uint32_t* __restrict__ segA, const uint32_t* __restrict__ segB, uint32_t t1,
                uint32_t* __restrict__ winres) {
    __shared__ Xyzz<F> sh[kReduceThreads];
    const int k = threadIdx.x;
    const uint32_t w = blockIdx.x;
    const uint32_t l2 = t1 >= kReduceThreads ? t1 / kReduceThreads : 1u;
    int log_l2 = 0;
    while ((1u << log_l2) < l2) ++log_l2;
    const size_t base = (size_t)w * t1;

    Xyzz<F> run = xyzz_inf<F>(), u = xyzz_inf<F>(), bs = xyzz_inf<F>();
    for (int l = (int)l2 - 1; l >= 0; --l) {
        const uint32_t t = (uint32_t)k * l2 + (uint32_t)l;
        if (t < t1) {
            run = xyzz_add_ol<F>(run, xyzz_load<F>(segA, base + t));
            u = xyzz_add_ol<F>(u, run);
            bs = xyzz_add_ol<F>(bs, xyzz_load<F>(segB, base + t));
        }
    }
    Xyzz<F> R = run;
    sh[k] = R;
    __syncthreads();
    for (int s = 1; s < kReduceThreads; s <<= 1) {
        Xyzz<F> other = (k + s < kReduceThreads) ? sh[k + s] : xyzz_inf<F>();
        __syncthreads();
        R = xyzz_add_ol<F>(R, other);
        sh[k] = R;
        __syncthreads();
    }
    const Xyzz<F> V = sh[0];
    __syncthreads();
    Xyzz<F> term = xyzz_add_ol<F>(bs, dbl_n<F>(xyzz_add_ol<F>(u, dbl_n<F>(R, log_l2)), 4));
    sh[k] = term;
    __syncthreads();
    for (int s = kReduceThreads / 2; s > 0; s >>= 1) {
        if (k < s) sh[k] = xyzz_add_ol<F>(sh[k], sh[k + s]);
        __syncthreads();
    }
    if (k == 0) {
        Xyzz<F> x = xyzz_add_ol<F>(dbl_n<F>(V, log_l2), V);
        x = dbl_n<F>(x, 4);
        xyzz_store<F>(winres, w, xyzz_add_ol<F>(sh[0], xyzz_neg<F>(x)));
    }
}

// --------------------------------------------------------- combine + output
template <class C>
__device__ void write_compressed(const Fe<typename C::Fp>& x_canon, const Fe<typename C::Fp>& y_canon,
                                 uint8_t* __restrict__ comp) {
    using F = typename C::Fp;
    const Fe<F> ny = fe_neg<F>(y_canon);
    bool larger = false;
    for (int k = F::N - 1; k >= 0; --k) {
        if (y_canon.v[k] != ny.v[k]) {
            larger = y_canon.v[k] > ny.v[k];
            break;
        }
    }
    fe_to_be_bytes<F>(x_canon, comp, C::FB);
    if (C::ID == 2) comp[0] |= 0x80 | (larger ? 0x20 : 0);
    else comp[0] |= larger ? 0xC0 : 0x80;
}

template <class C>
__device__ void write_infinity(uint8_t* __restrict__ out) {
    for (int k = 0; k < 3 * C::FB; ++k) out[k] = 0;
    out[2 * C::FB] = (C::ID == 2) ? 0xC0 : 0x40;
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

template <class C>
__global__ void k_combine(const uint32_t* __restrict__ winres, int nwin, int c, uint8_t* __restrict__ out) {
    using F = typename C::Fp;
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    Xyzz<F> r = xyzz_load<F>(winres, nwin - 1);
    for (int w = nwin - 2; w >= 0; --w) {
        if (!xyzz_is_inf<F>(r)) r = dbl_n<F>(r, c);
        r = xyzz_add_ol<F>(r, xyzz_load<F>(winres, w));
    }
    write_output<C>(r, out);
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

// ------------------------------------------------------------- host side
template <class C>
void msm_device_t(const uint32_t* d_pts, const uint8_t* d_inf, const uint32_t* d_scalars, size_t n,
                  const MsmParams& prm, MsmWorkspace& ws, hipStream_t s, uint8_t* d_out) {
    using F = typename C::Fp;
    const int c = prm.c ? prm.c : msm_default_window_bits(n);
    const int W = msm_num_windows(c);
    const uint32_t B = 1u << (c - 1);
    const uint32_t NB = (uint32_t)W * B;
    const uint32_t chunk = (uint32_t)(prm.chunk > 0 ? prm.chunk : 32);
    const size_t max_entries = n * (size_t)W;
    if (max_entries >= 0x80000000ull) throw std::invalid_argument("MSM too large for 31-bit entry indices");
    const uint32_t max_chunks = ceil_div(max_entries, chunk);
    const size_t pw = 4 * F::N * sizeof(uint32_t);  // XYZZ point bytes
    const uint32_t t1 = B / kSeg;

    ws.counts.ensure(NB * sizeof(uint32_t));
    ws.offsets.ensure((NB + 1) * sizeof(uint32_t));
    ws.cursor.ensure(NB * sizeof(uint32_t));
    ws.block_sums.ensure(scan_block_count(NB) * sizeof(uint32_t));
    ws.total.ensure(sizeof(uint32_t));
    ws.entries.ensure((max_entries + 1) * sizeof(uint32_t));
    ws.ebucket.ensure((max_entries + 1) * sizeof(uint32_t));
    ws.bsum.ensure(NB * pw);
    ws.carry.ensure((size_t)(max_chunks + 1) * pw);
    ws.segA.ensure((size_t)W * t1 * pw);
    ws.segB.ensure((size_t)W * t1 * pw);
    ws.winres.ensure((size_t)W * pw);

    uint32_t* counts = ws.counts.as<uint32_t>();
    KZ_HIP(hipMemsetAsync(counts, 0, NB * sizeof(uint32_t), s));
    digits_dispatch(c, false, d_scalars, d_inf, (uint32_t)n, counts, nullptr, nullptr, s);
    device_exclusive_scan(counts, ws.offsets.as<uint32_t>(), ws.cursor.as<uint32_t>(), NB,
                          ws.block_sums.as<uint32_t>(), ws.total.as<uint32_t>(), s);
    digits_dispatch(c, true, d_scalars, d_inf, (uint32_t)n, ws.cursor.as<uint32_t>(), ws.entries.as<uint32_t>(),
                    ws.ebucket.as<uint32_t>(), s);
    hipLaunchKernelGGL(k_accumulate<F>, dim3(ceil_div(max_chunks, 64)), dim3(64), 0, s,
                       ws.entries.as<uint32_t>(), ws.ebucket.as<uint32_t>(), ws.offsets.as<uint32_t>(),
                       ws.total.as<uint32_t>(), chunk, d_pts, ws.bsum.as<uint32_t>(), ws.carry.as<uint32_t>());
    hipLaunchKernelGGL(k_fixup<F>, dim3(ceil_div(NB, 64)), dim3(64), 0, s, counts, ws.offsets.as<uint32_t>(),
                       chunk, NB, ws.bsum.as<uint32_t>(), ws.carry.as<uint32_t>());
    const uint32_t nseg = (uint32_t)W * t1;
    hipLaunchKernelGGL(k_reduce_seg<F>, dim3(ceil_div(nseg, 64)), dim3(64), 0, s, ws.bsum.as<uint32_t>(), nseg,
                       ws.segA.as<uint32_t>(), ws.segB.as<uint32_t>());
    hipLaunchKernelGGL(k_reduce_window<F>, dim3(W), dim3(kReduceThreads), 0, s, ws.segA.as<uint32_t>(),
                       ws.segB.as<uint32_t>(), t1, ws.winres.as<uint32_t>());
    hipLaunchKernelGGL(k_combine<C>, dim3(1), dim3(64), 0, s, ws.winres.as<uint32_t>(), W, c, d_out);
    KZ_HIP(hipGetLastError());
}

template <class C>
void points_from_wire_t(const uint8_t* d_wire, uint32_t* d_pts, uint8_t* d_inf, size_t n, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_points_from_wire<C>, dim3(ceil_div(n, 128)), dim3(128), 0, s, d_wire, d_pts, d_inf,
                       (uint32_t)n);
    KZ_HIP(hipGetLastError());
}

template <class C>
void compress_t(const uint8_t* d_wire, uint8_t* d_out, hipStream_t s) {
    hipLaunchKernelGGL(k_compress_point<C>, dim3(1), dim3(64), 0, s, d_wire, d_out);
    KZ_HIP(hipGetLastError());
}

}  // namespace kz
