// Signed-digit decomposition and counting sort of (window, bucket) keys for
// the Pippenger MSM; curve independent.
//
// Digits follow the reference (gpukit/cpp/cpu_reference/multi_pippenger.cpp:90-134):
// the c-bit field of the 256-bit LE scalar at bit w*c plus the carry from the
// previous window; a value above 2^(c-1) becomes value - 2^c with a carry of 1.
// Digit d != 0 goes to bucket |d| - 1 of window w, negated when d < 0.  Here the
// windows continue until the carry is absorbed (one extra window when c
// divides 256), where the reference drops it.
//
// Keys are global bucket ids gb = w * 2^(c-1) + (|d| - 1).  One pass counts
// them (atomic histogram), a device scan turns counts into offsets, and a
// second pass scatters (point index | sign << 31) to its bucket's slot.  The
// order inside a bucket is whatever the atomics give; bucket sums do not
// depend on it (the group is commutative), so results stay bit-exact.
//
// Window-table MSMs (table_stride = T > 0): the points are a table whose row w
// holds 2^(c w) P_i at index w T + i, so digit d of window w is the point
// w T + i in bucket |d| - 1 of a single window shared by all rows: keys are
// |d| - 1 and entries (w T + i | sign << 31).
#include "msm_internal.h"
#include "scan.h"

namespace kz {

// Windows of c-bit signed digits over BITS-bit scalars (256, or 127 / 128 for
// the magnitudes of a signed GLV split: BN254 and BLS12-381 / secp256k1),
// with the extra window for the final carry when c divides BITS (never for
// 127).
template <int CB, int BITS = 256>
struct Digits {
    static constexpr int W0 = (BITS + CB - 1) / CB;
    static constexpr int W = W0 + ((W0 * CB == BITS) ? 1 : 0);
    static constexpr uint32_t B = 1u << (CB - 1);
};

template <int CB>
__device__ __forceinline__ uint32_t window_bits(const uint32_t (&s)[8], int w) {
    const int off = w * CB;
    const int lo = off >> 5;
    const int sh = off & 31;
    uint32_t v = 0;
    if (lo < 8) {
        v = s[lo] >> sh;
        if (sh + CB > 32 && lo + 1 < 8) v |= s[lo + 1] << (32 - sh);
    }
    return v & ((1u << CB) - 1u);
}

__device__ __forceinline__ void load_scalar(const uint32_t* __restrict__ scalars, uint32_t i, uint32_t (&s)[8]) {
    const uint4* q = reinterpret_cast<const uint4*>(scalars + (size_t)i * 8);
    uint4 a = q[0], b = q[1];
    s[0] = a.x; s[1] = a.y; s[2] = a.z; s[3] = a.w;
    s[4] = b.x; s[5] = b.y; s[6] = b.z; s[7] = b.w;
}

// Scalars of at most BITS bits: 32-byte records, or for at most 128 bits
// (the magnitudes of a GLV split) 16-byte records.
template <int BITS>
__device__ __forceinline__ void load_scalar_bits(const uint32_t* __restrict__ scalars, uint32_t i, uint32_t (&s)[8]) {
    if (BITS > 128) {
        load_scalar(scalars, i, s);
    } else {
        const uint4 a = reinterpret_cast<const uint4*>(scalars)[i];
        s[0] = a.x; s[1] = a.y; s[2] = a.z; s[3] = a.w;
        s[4] = s[5] = s[6] = s[7] = 0;
    }
}

template <int CB, bool SCATTER>
__global__ void __launch_bounds__(256)
k_digits(const uint32_t* __restrict__ scalars, const uint8_t* __restrict__ inf, uint32_t n, uint32_t table_stride,
         uint32_t* __restrict__ counts_or_cursor, uint32_t* __restrict__ entries) {
    using D = Digits<CB>;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (inf && inf[i]) return;
    uint32_t s[8];
    load_scalar(scalars, i, s);
    uint32_t carry = 0;
#pragma unroll
    for (int w = 0; w < D::W; ++w) {
        int32_t d = (int32_t)(window_bits<CB>(s, w) + carry);
        carry = 0;
        if (d > (int32_t)D::B) {
            d -= (int32_t)(1u << CB);
            carry = 1;
        }
        if (d != 0) {
            const uint32_t gb = (table_stride ? 0u : (uint32_t)w * D::B) + (uint32_t)(d > 0 ? d : -d) - 1u;
            if (!SCATTER) {
                atomicAdd(&counts_or_cursor[gb], 1u);
            } else {
                const uint32_t pos = atomicAdd(&counts_or_cursor[gb], 1u);
                entries[pos] = ((uint32_t)w * table_stride + i) | (d < 0 ? 0x80000000u : 0u);
            }
        }
    }
}

template <int CB>
static void launch(bool scatter, const uint32_t* d_scalars, const uint8_t* d_inf, uint32_t n, uint32_t table_stride,
                   uint32_t* counts, uint32_t* entries, hipStream_t s) {
    const unsigned grid = ceil_div(n, 256);
    if (!scatter)
        hipLaunchKernelGGL((k_digits<CB, false>), dim3(grid), dim3(256), 0, s, d_scalars, d_inf, n, table_stride,
                           counts, nullptr);
    else
        hipLaunchKernelGGL((k_digits<CB, true>), dim3(grid), dim3(256), 0, s, d_scalars, d_inf, n, table_stride,
                           counts, entries);
}

// ------------------------------------------------------------- LDS sort
// Single MSMs with c <= 16 sort without global atomics.  Global atomics from
// all eight XCDs meet at the memory side (each XCD has its own L2), which made
// the atomic count and scatter passes the slowest part of the sort.  Here a
// workgroup owns one (window w, part p) pair -- part p = a contiguous range of
// points -- and counts its digits in an LDS histogram of 2^(c-1) words
// (128 KiB at c = 16; gfx950 gives a workgroup up to 160 KiB):
//
//   k_sort_count    hist[b] -> cnt[w][p][b]
//   k_sort_colscan  pre[w][p][b] = sum_{p' < p} cnt[w][p'][b],
//                   tot[w][b] = sum_p cnt[w][p][b]
//   scan            offsets[w B + b] = exclusive prefix of tot
//   k_sort_scatter  LDS cursor[b] = offsets[w B + b] + pre[w][p][b];
//                   each digit takes a slot by an LDS atomic
//
// Workgroup ids interleave windows (id = p * wpad + w, wpad a multiple of 8),
// so the workgroups of one window run on one XCD (ids go round-robin over
// the eight XCDs) and its scattered entry writes meet in that XCD's L2.
constexpr int kSortThreads = 1024;

// Signed digit of window w (the carry chain runs over the lower windows).
template <int CB, int BITS>
__device__ __forceinline__ int32_t digit_of_window(const uint32_t (&s)[8], uint32_t w) {
    using D = Digits<CB, BITS>;
    uint32_t carry = 0;
    int32_t sel = 0;
#pragma unroll
    for (int ww = 0; ww < D::W; ++ww) {
        int32_t d = (int32_t)(window_bits<CB>(s, ww) + carry);
        carry = 0;
        if (d > (int32_t)D::B) {
            d -= (int32_t)(1u << CB);
            carry = 1;
        }
        if ((uint32_t)ww == w) sel = d;
    }
    return sel;
}

template <int CB, int BITS>
__global__ void __launch_bounds__(kSortThreads)
k_sort_count(const uint32_t* __restrict__ scalars, const uint8_t* __restrict__ inf, uint32_t n, uint32_t per_part,
             uint32_t wpad, uint32_t nparts, uint32_t* __restrict__ cnt) {
    using D = Digits<CB, BITS>;
    __shared__ uint32_t hist[D::B];
    const uint32_t w = blockIdx.x % wpad, p = blockIdx.x / wpad;
    if (w >= (uint32_t)D::W) return;  // padding workgroup (whole group leaves)
    for (uint32_t b = threadIdx.x; b < D::B; b += kSortThreads) hist[b] = 0;
    __syncthreads();
    const uint32_t lo = p * per_part, hi = min(n, lo + per_part);
    for (uint32_t i = lo + threadIdx.x; i < hi; i += kSortThreads) {
        if (inf && inf[i]) continue;
        uint32_t s[8];
        load_scalar_bits<BITS>(scalars, i, s);
        const int32_t d = digit_of_window<CB, BITS>(s, w);
        if (d != 0) atomicAdd(&hist[(d > 0 ? d : -d) - 1], 1u);
    }
    __syncthreads();
    uint32_t* out = cnt + ((size_t)w * nparts + p) * D::B;
    for (uint32_t b = threadIdx.x; b < D::B; b += kSortThreads) out[b] = hist[b];
}

// One thread per (window, bucket): prefix over the parts.
__global__ void __launch_bounds__(256)
k_sort_colscan(const uint32_t* __restrict__ cnt, uint32_t* __restrict__ pre, uint32_t* __restrict__ tot,
               uint32_t nwin, uint32_t nparts, uint32_t nbw) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nwin * nbw) return;
    const uint32_t w = t / nbw, b = t - w * nbw;
    uint32_t run = 0;
    for (uint32_t p = 0; p < nparts; ++p) {
        const size_t idx = ((size_t)w * nparts + p) * nbw + b;
        const uint32_t c = cnt[idx];
        pre[idx] = run;
        run += c;
    }
    tot[t] = run;
}

template <int CB, int BITS>
__global__ void __launch_bounds__(kSortThreads)
k_sort_scatter(const uint32_t* __restrict__ scalars, const uint8_t* __restrict__ inf, uint32_t n,
               uint32_t per_part, uint32_t wpad, uint32_t nparts, uint32_t table_stride,
               const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ pre,
               uint32_t* __restrict__ entries) {
    using D = Digits<CB, BITS>;
    __shared__ uint32_t cur[D::B];
    const uint32_t w = blockIdx.x % wpad, p = blockIdx.x / wpad;
    if (w >= (uint32_t)D::W) return;
    const uint32_t* pp = pre + ((size_t)w * nparts + p) * D::B;
    const uint32_t* wo = offsets + (table_stride ? 0u : w * D::B);
    const uint32_t base = w * table_stride;
    for (uint32_t b = threadIdx.x; b < D::B; b += kSortThreads) cur[b] = wo[b] + pp[b];
    __syncthreads();
    const uint32_t lo = p * per_part, hi = min(n, lo + per_part);
    for (uint32_t i = lo + threadIdx.x; i < hi; i += kSortThreads) {
        if (inf && inf[i]) continue;
        uint32_t s[8];
        load_scalar_bits<BITS>(scalars, i, s);
        const int32_t d = digit_of_window<CB, BITS>(s, w);
        if (d != 0) {
            const uint32_t pos = atomicAdd(&cur[(d > 0 ? d : -d) - 1], 1u);
            entries[pos] = (base + i) | (d < 0 ? 0x80000000u : 0u);
        }
    }
}

bool lds_sort_supported(int c, size_t n) { return c >= 5 && c <= 16 && n >= 4096; }

void lds_sort_shape(int c, int bits, size_t n, uint32_t& wpad, uint32_t& nparts, uint32_t& per_part) {
    const uint32_t W = (uint32_t)msm_num_windows(c, bits);
    wpad = (W + 7) / 8 * 8;
    // parts of >= 4096 points, doubled while the launch stays within 256
    // (window, part) workgroups: one per CU of the MI355X (a 16-bit window's
    // histogram takes 128 KiB of LDS, so one workgroup fills a CU).  Fewer,
    // longer parts write each bucket's entries in longer runs: at 2^21 GLV
    // points and 16-bit windows (9 windows then), 64 parts per window made
    // the scattered 4-byte entry writes cost about 7x their bytes in HBM
    // writes and the sort took 0.63 ms; 16 parts took 0.46 ms, 8 parts
    // 0.73 ms.  Half the CUs idle also cost: 16 parts of 8 windows at 2^23
    // points sorted in 2.96 ms against 2.57 ms for 16 parts of 9.
    uint32_t parts = 1;
    while (parts * 2 * W <= 256 && ((size_t)parts * 2) * 4096 <= n) parts *= 2;
    nparts = parts;
    per_part = (uint32_t)((n + parts - 1) / parts);
}

template <int CB, int BITS>
static void lds_sort_launch(const uint32_t* d_scalars, const uint8_t* d_inf, uint32_t n, uint32_t table_stride,
                            uint32_t* cnt, uint32_t* pre, uint32_t* tot, uint32_t* offsets, uint32_t* block_sums,
                            uint32_t* total, uint32_t* entries, hipStream_t s) {
    using D = Digits<CB, BITS>;
    uint32_t wpad, nparts, per_part;
    lds_sort_shape(CB, BITS, n, wpad, nparts, per_part);
    const unsigned grid = wpad * nparts;
    hipLaunchKernelGGL((k_sort_count<CB, BITS>), dim3(grid), dim3(kSortThreads), 0, s, d_scalars, d_inf, n, per_part, wpad,
                       nparts, cnt);
    // per-window buckets: prefix over the parts of each window; one shared
    // bucket set (window table): the (window, part) rows are the parts
    const uint32_t nwin = table_stride ? 1u : (uint32_t)D::W;
    const uint32_t rows = table_stride ? (uint32_t)D::W * nparts : nparts;
    const uint32_t nb = nwin * D::B;
    hipLaunchKernelGGL(k_sort_colscan, dim3(ceil_div(nb, 256)), dim3(256), 0, s, cnt, pre, tot, nwin, rows, D::B);
    device_exclusive_scan(tot, offsets, nullptr, nb, block_sums, total, s);
    hipLaunchKernelGGL((k_sort_scatter<CB, BITS>), dim3(grid), dim3(kSortThreads), 0, s, d_scalars, d_inf, n, per_part, wpad,
                       nparts, table_stride, offsets, pre, entries);
}

void lds_sort_dispatch(int c, int bits, const uint32_t* d_scalars, const uint8_t* d_inf, uint32_t n,
                       uint32_t table_stride, uint32_t* cnt, uint32_t* pre, uint32_t* tot, uint32_t* offsets,
                       uint32_t* block_sums, uint32_t* total, uint32_t* entries, hipStream_t s) {
    if (bits != 256 && bits != 127 && bits != 128)
        throw std::invalid_argument("LDS sort: scalar bits must be 256, 127 or 128");
    switch (c) {
#define KZ_C(CB)                                                                                                     \
    case CB:                                                                                                         \
        if (bits == 256)                                                                                             \
            lds_sort_launch<CB, 256>(d_scalars, d_inf, n, table_stride, cnt, pre, tot, offsets, block_sums, total,  \
                                     entries, s);                                                                    \
        else if (bits == 127)                                                                                        \
            lds_sort_launch<CB, 127>(d_scalars, d_inf, n, table_stride, cnt, pre, tot, offsets, block_sums, total,  \
                                     entries, s);                                                                    \
        else                                                                                                         \
            lds_sort_launch<CB, 128>(d_scalars, d_inf, n, table_stride, cnt, pre, tot, offsets, block_sums, total,  \
                                     entries, s);                                                                    \
        break;
        KZ_C(5) KZ_C(6) KZ_C(7) KZ_C(8) KZ_C(9) KZ_C(10) KZ_C(11) KZ_C(12) KZ_C(13) KZ_C(14) KZ_C(15) KZ_C(16)
#undef KZ_C
        default: throw std::invalid_argument("LDS sort: window bits out of range [5, 16]");
    }
}

// Batch form: the window size is a per-job runtime value, so the windows are
// read by streaming the scalar's words through a 64-bit bit buffer (static
// word indices only: no scratch).  Block b serves job j = last job with
// blk_base <= b; its global bucket ids start at the job's gb_base and its
// entries carry the point's index in the batch's point array.  A window-table
// job (jd.table = T > 0) sends the digit of every window w to its one set of
// buckets, as the entry of point tab_base + pt_base + w T + i.
template <bool SCATTER>
__global__ void __launch_bounds__(256)
k_digits_multi(const uint32_t* __restrict__ scalars, const uint8_t* __restrict__ inf,
               const MsmJob* __restrict__ jobs, uint32_t njobs, uint32_t* __restrict__ counts_or_cursor,
               uint32_t* __restrict__ entries) {
    __shared__ MsmJob sj;
    if (threadIdx.x == 0) {
        uint32_t lo = 0, hi = njobs;  // invariant: jobs[lo].blk_base <= blockIdx.x
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (jobs[mid].blk_base <= blockIdx.x) lo = mid;
            else hi = mid;
        }
        sj = jobs[lo];
    }
    __syncthreads();
    const MsmJob jd = sj;
    const uint32_t i = (blockIdx.x - jd.blk_base) * blockDim.x + threadIdx.x;
    if (i >= jd.n) return;
    const uint32_t p = jd.pt_base + i;
    if (inf && inf[p]) return;
    uint32_t s[8];
    load_scalar(scalars + jd.sc_base * 8, i, s);
    const uint32_t c = jd.c, B = 1u << (c - 1), mask = (1u << c) - 1u;
    const uint32_t w0 = (256u + c - 1u) / c, wd = w0 + (w0 * c == 256u ? 1u : 0u);  // digit windows
    const uint32_t wb = jd.table ? 0u : B;  // bucket-id step per digit window
    uint32_t w = 0, carry = 0;
    auto emit = [&](uint32_t v) {
        int32_t d = (int32_t)(v + carry);
        carry = 0;
        if (d > (int32_t)B) {
            d -= (int32_t)(1u << c);
            carry = 1;
        }
        if (d != 0) {
            const uint32_t gb = jd.gb_base + w * wb + (uint32_t)(d > 0 ? d : -d) - 1u;
            if (!SCATTER) {
                atomicAdd(&counts_or_cursor[gb], 1u);
            } else {
                const uint32_t pos = atomicAdd(&counts_or_cursor[gb], 1u);
                entries[pos] = (jd.tab_base + p + w * jd.table) | (d < 0 ? 0x80000000u : 0u);
            }
        }
        ++w;
    };
    uint64_t buf = 0;
    uint32_t nb = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        buf |= (uint64_t)s[k] << nb;
        nb += 32;
        while (nb >= c) {
            emit((uint32_t)buf & mask);
            buf >>= c;
            nb -= c;
        }
    }
    while (w < wd) {  // the remaining < c bits, then the carry window
        emit((uint32_t)buf & mask);
        buf >>= c;
    }
}

void digits_multi_dispatch(bool scatter, const uint32_t* d_scalars, const uint8_t* d_inf, const MsmJob* d_jobs,
                           uint32_t njobs, uint32_t nblocks, uint32_t* counts_or_cursor, uint32_t* entries,
                           hipStream_t s) {
    if (nblocks == 0) return;
    if (!scatter)
        hipLaunchKernelGGL(k_digits_multi<false>, dim3(nblocks), dim3(256), 0, s, d_scalars, d_inf, d_jobs, njobs,
                           counts_or_cursor, nullptr);
    else
        hipLaunchKernelGGL(k_digits_multi<true>, dim3(nblocks), dim3(256), 0, s, d_scalars, d_inf, d_jobs, njobs,
                           counts_or_cursor, entries);
}

void digits_dispatch(int c, bool scatter, const uint32_t* d_scalars, const uint8_t* d_inf, uint32_t n,
                     uint32_t table_stride, uint32_t* counts_or_cursor, uint32_t* entries, hipStream_t s) {
    switch (c) {
#define KZ_C(CB) \
    case CB: launch<CB>(scatter, d_scalars, d_inf, n, table_stride, counts_or_cursor, entries, s); break;
        KZ_C(5) KZ_C(6) KZ_C(7) KZ_C(8) KZ_C(9) KZ_C(10) KZ_C(11) KZ_C(12) KZ_C(13) KZ_C(14) KZ_C(15)
        KZ_C(16) KZ_C(17) KZ_C(18)
#undef KZ_C
        default: throw std::invalid_argument("window bits out of range [5, 18]");
    }
}

// ------------------------------------------------------------ batch LDS sort
// The global-atomic passes of k_digits_multi make the sort the slowest stage
// of a batch of small and mid-size MSMs (hundreds of millions of digits, all
// through memory-side atomics).  The batch form of the LDS sort:
//
//   k_batch_digits   one thread per point writes its signed digits as 16-bit
//                    codes, window-major per job (coalesced)
//   k_batch_count    one workgroup per (job, digit window w, part p): LDS
//                    histogram of 2^(c-1) words -> cnt[item]
//   k_batch_colscan  per global bucket: prefix over the job's parts (over all
//                    its (window, part) items for a window-table job, whose
//                    windows share one set of buckets) -> pre, and the
//                    bucket's count -> tot
//   scan             offsets
//   k_batch_scatter  per item: LDS cursor[b] = offsets[gb] + pre[item][b];
//                    every digit takes a slot by an LDS atomic
//
// Digit code: d > 0 -> d - 1; d < 0 -> 0x8000 | (|d| - 1); d = 0 -> 0xFFFF.
// (|d| <= 2^(c-1) <= 2^15 for positive digits and < 2^(c-1) for negative
// ones, so the three ranges do not meet.)
constexpr uint16_t kNoDigit = 0xFFFF;

__device__ __forceinline__ uint32_t find_job_by_block(const MsmJob* __restrict__ jobs, uint32_t njobs, uint32_t b) {
    uint32_t lo = 0, hi = njobs;  // invariant: jobs[lo].blk_base <= b
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (jobs[mid].blk_base <= b) lo = mid;
        else hi = mid;
    }
    return lo;
}

__global__ void __launch_bounds__(256)
k_batch_digits(const uint32_t* __restrict__ scalars, const uint8_t* __restrict__ inf, const MsmJob* __restrict__ jobs,
               const SortJob* __restrict__ sjobs, uint32_t njobs, uint16_t* __restrict__ dig) {
    __shared__ uint32_t sj;
    if (threadIdx.x == 0) sj = find_job_by_block(jobs, njobs, blockIdx.x);
    __syncthreads();
    const MsmJob jd = jobs[sj];
    const SortJob so = sjobs[sj];
    const uint32_t i = (blockIdx.x - jd.blk_base) * blockDim.x + threadIdx.x;
    if (i >= jd.n) return;
    uint16_t* out = dig + so.dig_base + i;
    const size_t stride = jd.n;
    if (inf && inf[jd.pt_base + i]) {
        for (uint32_t w = 0; w < so.wd; ++w) out[w * stride] = kNoDigit;
        return;
    }
    uint32_t s[8];
    load_scalar(scalars + jd.sc_base * 8, i, s);
    const uint32_t c = jd.c, B = 1u << (c - 1), mask = (1u << c) - 1u;
    uint32_t w = 0, carry = 0;
    auto emit = [&](uint32_t v) {
        int32_t d = (int32_t)(v + carry);
        carry = 0;
        if (d > (int32_t)B) {
            d -= (int32_t)(1u << c);
            carry = 1;
        }
        const uint16_t code = d > 0 ? (uint16_t)(d - 1) : d < 0 ? (uint16_t)(0x8000u | (uint32_t)(-d - 1)) : kNoDigit;
        out[w * stride] = code;
        ++w;
    };
    uint64_t buf = 0;
    uint32_t nbits = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        buf |= (uint64_t)s[k] << nbits;
        nbits += 32;
        while (nbits >= c) {
            emit((uint32_t)buf & mask);
            buf >>= c;
            nbits -= c;
        }
    }
    while (w < so.wd) {  // the remaining < c bits, then the carry window
        emit((uint32_t)buf & mask);
        buf >>= c;
    }
}

// Job of sort item `item` (workgroup), read by thread 0 into LDS.
__device__ __forceinline__ uint32_t find_job_by_item(const SortJob* __restrict__ sjobs, uint32_t njobs, uint32_t item) {
    uint32_t lo = 0, hi = njobs;  // invariant: sjobs[lo].item_base <= item
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (sjobs[mid].item_base <= item) lo = mid;
        else hi = mid;
    }
    return lo;
}

__global__ void __launch_bounds__(kSortThreads)
k_batch_count(const uint16_t* __restrict__ dig, const MsmJob* __restrict__ jobs, const SortJob* __restrict__ sjobs,
              uint32_t njobs, uint32_t* __restrict__ cnt) {
    extern __shared__ uint32_t hist[];
    __shared__ uint32_t sj;
    if (threadIdx.x == 0) sj = find_job_by_item(sjobs, njobs, blockIdx.x);
    __syncthreads();
    const MsmJob jd = jobs[sj];
    const SortJob so = sjobs[sj];
    const uint32_t item = blockIdx.x - so.item_base, w = item / so.parts, p = item - w * so.parts;
    const uint32_t B = 1u << (jd.c - 1);
    for (uint32_t b = threadIdx.x; b < B; b += kSortThreads) hist[b] = 0;
    __syncthreads();
    const uint16_t* d = dig + so.dig_base + (size_t)w * jd.n;
    const uint32_t lo = p * so.per_part, hi = min(jd.n, lo + so.per_part);
    for (uint32_t i = lo + threadIdx.x; i < hi; i += kSortThreads) {
        const uint16_t code = d[i];
        if (code != kNoDigit) atomicAdd(&hist[code & 0x7FFFu], 1u);
    }
    __syncthreads();
    uint32_t* out = cnt + so.cnt_base + (size_t)item * B;
    for (uint32_t b = threadIdx.x; b < B; b += kSortThreads) out[b] = hist[b];
}

// One thread per global bucket gb = gb_base + w B + b of job j.
__global__ void __launch_bounds__(256)
k_batch_colscan(const uint32_t* __restrict__ cnt, const MsmJob* __restrict__ jobs, const SortJob* __restrict__ sjobs,
                uint32_t njobs, uint32_t nb, uint32_t* __restrict__ pre, uint32_t* __restrict__ tot) {
    const uint32_t gb = blockIdx.x * blockDim.x + threadIdx.x;
    if (gb >= nb) return;
    uint32_t lo = 0, hi = njobs;  // invariant: jobs[lo].gb_base <= gb
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (jobs[mid].gb_base <= gb) lo = mid;
        else hi = mid;
    }
    const MsmJob jd = jobs[lo];
    const SortJob so = sjobs[lo];
    const uint32_t B = 1u << (jd.c - 1), r = gb - jd.gb_base, b = r & (B - 1);
    // items (w, p) feeding this bucket: one window's parts, or every item
    const uint32_t first = jd.table ? 0u : (r >> (jd.c - 1)) * so.parts;
    const uint32_t last = jd.table ? so.wd * so.parts : first + so.parts;
    uint32_t run = 0;
    for (uint32_t it = first; it < last; ++it) {
        const size_t idx = so.cnt_base + (size_t)it * B + b;
        const uint32_t c = cnt[idx];
        pre[idx] = run;
        run += c;
    }
    tot[gb] = run;
}

__global__ void __launch_bounds__(kSortThreads)
k_batch_scatter(const uint16_t* __restrict__ dig, const MsmJob* __restrict__ jobs, const SortJob* __restrict__ sjobs,
                uint32_t njobs, const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ pre,
                uint32_t* __restrict__ entries) {
    extern __shared__ uint32_t cur[];
    __shared__ uint32_t sj;
    if (threadIdx.x == 0) sj = find_job_by_item(sjobs, njobs, blockIdx.x);
    __syncthreads();
    const MsmJob jd = jobs[sj];
    const SortJob so = sjobs[sj];
    const uint32_t item = blockIdx.x - so.item_base, w = item / so.parts, p = item - w * so.parts;
    const uint32_t B = 1u << (jd.c - 1);
    const uint32_t* pp = pre + so.cnt_base + (size_t)item * B;
    const uint32_t* wo = offsets + jd.gb_base + (jd.table ? 0u : w * B);
    const uint32_t base = jd.tab_base + jd.pt_base + w * jd.table;
    for (uint32_t b = threadIdx.x; b < B; b += kSortThreads) cur[b] = wo[b] + pp[b];
    __syncthreads();
    const uint16_t* d = dig + so.dig_base + (size_t)w * jd.n;
    const uint32_t lo = p * so.per_part, hi = min(jd.n, lo + so.per_part);
    for (uint32_t i = lo + threadIdx.x; i < hi; i += kSortThreads) {
        const uint16_t code = d[i];
        if (code != kNoDigit) {
            const uint32_t pos = atomicAdd(&cur[code & 0x7FFFu], 1u);
            entries[pos] = (base + i) | ((uint32_t)(code & 0x8000u) << 16);
        }
    }
}

bool batch_lds_sort_prepare(size_t dyn_lds) {
    if (dyn_lds > 160 * 1024) return false;
    if (dyn_lds <= 64 * 1024) return true;
    return hipFuncSetAttribute(reinterpret_cast<const void*>(&k_batch_count),
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn_lds) == hipSuccess &&
           hipFuncSetAttribute(reinterpret_cast<const void*>(&k_batch_scatter),
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn_lds) == hipSuccess;
}

void batch_lds_sort_dispatch(const uint32_t* d_scalars, const uint8_t* d_inf, const MsmJob* d_jobs,
                             const SortJob* d_sjobs, uint32_t njobs, uint32_t dig_blocks, uint32_t items, uint32_t nb,
                             size_t dyn_lds, uint16_t* dig, uint32_t* cnt, uint32_t* pre, uint32_t* tot,
                             uint32_t* offsets, uint32_t* block_sums, uint32_t* total, uint32_t* entries,
                             hipStream_t s) {
    if (dig_blocks == 0 || items == 0) return;
    hipLaunchKernelGGL(k_batch_digits, dim3(dig_blocks), dim3(256), 0, s, d_scalars, d_inf, d_jobs, d_sjobs, njobs,
                       dig);
    hipLaunchKernelGGL(k_batch_count, dim3(items), dim3(kSortThreads), dyn_lds, s, dig, d_jobs, d_sjobs, njobs, cnt);
    hipLaunchKernelGGL(k_batch_colscan, dim3(ceil_div(nb, 256)), dim3(256), 0, s, cnt, d_jobs, d_sjobs, njobs, nb,
                       pre, tot);
    device_exclusive_scan(tot, offsets, nullptr, nb, block_sums, total, s);
    hipLaunchKernelGGL(k_batch_scatter, dim3(items), dim3(kSortThreads), dyn_lds, s, dig, d_jobs, d_sjobs, njobs,
                       offsets, pre, entries);
}

}  // namespace kz
