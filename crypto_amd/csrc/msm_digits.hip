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
#include "msm_internal.h"

namespace kz {

template <int CB>
struct Digits {
    static constexpr int W0 = (256 + CB - 1) / CB;
    static constexpr int W = W0 + ((W0 * CB == 256) ? 1 : 0);
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

template <int CB, bool SCATTER>
__global__ void __launch_bounds__(256)
k_digits(const uint32_t* __restrict__ scalars, const uint8_t* __restrict__ inf, uint32_t n,
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
            const uint32_t gb = (uint32_t)w * D::B + (uint32_t)(d > 0 ? d : -d) - 1u;
            if (!SCATTER) {
                atomicAdd(&counts_or_cursor[gb], 1u);
            } else {
                const uint32_t pos = atomicAdd(&counts_or_cursor[gb], 1u);
                entries[pos] = i | (d < 0 ? 0x80000000u : 0u);
            }
        }
    }
}

template <int CB>
static void launch(bool scatter, const uint32_t* d_scalars, const uint8_t* d_inf, uint32_t n, uint32_t* counts,
                   uint32_t* entries, hipStream_t s) {
    const unsigned grid = ceil_div(n, 256);
    if (!scatter)
        hipLaunchKernelGGL((k_digits<CB, false>), dim3(grid), dim3(256), 0, s, d_scalars, d_inf, n, counts,
                           nullptr);
    else
        hipLaunchKernelGGL((k_digits<CB, true>), dim3(grid), dim3(256), 0, s, d_scalars, d_inf, n, counts,
                           entries);
}

// Batch form: the window size is a per-job runtime value, so the windows are
// read by streaming the scalar's words through a 64-bit bit buffer (static
// word indices only: no scratch).  Block b serves job j = last job with
// blk_base <= b; its global bucket ids start at the job's gb_base and its
// entries carry the point's index in the batch's point array.
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
    uint32_t w = 0, carry = 0;
    auto emit = [&](uint32_t v) {
        int32_t d = (int32_t)(v + carry);
        carry = 0;
        if (d > (int32_t)B) {
            d -= (int32_t)(1u << c);
            carry = 1;
        }
        if (d != 0) {
            const uint32_t gb = jd.gb_base + w * B + (uint32_t)(d > 0 ? d : -d) - 1u;
            if (!SCATTER) {
                atomicAdd(&counts_or_cursor[gb], 1u);
            } else {
                const uint32_t pos = atomicAdd(&counts_or_cursor[gb], 1u);
                entries[pos] = p | (d < 0 ? 0x80000000u : 0u);
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
    while (w < jd.nwin) {  // the remaining < c bits, then the carry window
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
                     uint32_t* counts_or_cursor, uint32_t* entries, hipStream_t s) {
    switch (c) {
#define KZ_C(CB) \
    case CB: launch<CB>(scatter, d_scalars, d_inf, n, counts_or_cursor, entries, s); break;
        KZ_C(5) KZ_C(6) KZ_C(7) KZ_C(8) KZ_C(9) KZ_C(10) KZ_C(11) KZ_C(12) KZ_C(13) KZ_C(14) KZ_C(15)
        KZ_C(16) KZ_C(17) KZ_C(18)
#undef KZ_C
        default: throw std::invalid_argument("window bits out of range [5, 18]");
    }
}

}  // namespace kz
