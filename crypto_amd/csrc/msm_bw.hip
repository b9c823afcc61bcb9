// Banderwagon (curve 3): the MSM stages of msm_impl.h over the twisted Edwards
// point operations (point_ops.h selects them for BwFp), and the wire reader,
// which validates like the reference's Element::deserialize_uncompressed
// (banderwagon/cpp/element.cpp:330-355) and Fr::from_bytes_le (fr.cpp:305-328).
#include "msm_impl.h"
#include "msm_stages.h"
#include "scan.h"

namespace kz {

using Bw = CurveBanderwagon;
using BwF = Bw::Fp;

// Wire (64 bytes: BE x || y) -> affine record (x, y, d x y) in Montgomery form.
// All-zero bytes are the identity (flagged, skipped by the digit sort).  A
// point is accepted when x and y are below q, d x^2 - 1 != 0,
// y^2 (d x^2 - 1) = a x^2 - 1 (y^2 = (a x^2 - 1) / (d x^2 - 1)) and 1 - a x^2
// is a non-zero square (the subgroup test); otherwise bit 0 of *err is set.
__global__ void __launch_bounds__(128)
k_bw_points_from_wire(const uint8_t* __restrict__ wire, uint32_t* __restrict__ pts, uint8_t* __restrict__ inf,
                      uint32_t n, uint32_t* __restrict__ err) {
    using F = BwF;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t* w = reinterpret_cast<const uint32_t*>(wire + (size_t)i * 64);
    uint32_t xw[8], yw[8], any = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const uint32_t wx = w[7 - k], wy = w[15 - k];
        xw[k] = __builtin_bswap32(wx);
        yw[k] = __builtin_bswap32(wy);
        any |= wx | wy;
    }
    EdAff<F> a;
    a.x = fe_zero<F>();
    a.y = fe_one<F>();
    a.t = fe_zero<F>();
    if (!any) {
        ed_aff_store<F>(pts, i, a);
        inf[i] = 1;
        return;
    }
    const Fe<F> x = fe_from_words<F>(xw, 8), y = fe_from_words<F>(yw, 8);
    bool ok = fe_is_canonical<F>(x) && fe_is_canonical<F>(y);
    if (ok) {
        a.x = fe_to_mont<F>(x);
        a.y = fe_to_mont<F>(y);
        const Fe<F> x2 = fe_mul_call<F>(a.x, a.x);
        const Fe<F> five_x2 = fe_mul5<F>(x2);
        const Fe<F> num = fe_sub<F>(fe_neg<F>(five_x2), fe_one<F>());  // a x^2 - 1
        const Fe<F> den = fe_sub<F>(fe_mul_call<F>(x2, fe_const<F>(EdConst::D)), fe_one<F>());
        ok = !fe_is_zero<F>(den) && fe_eq<F>(fe_mul_call<F>(fe_mul_call<F>(a.y, a.y), den), num);
        ok = ok && fe_is_square_nonzero<F>(fe_add<F>(fe_one<F>(), five_x2));  // 1 - a x^2
        a.t = fe_mul_call<F>(fe_mul_call<F>(a.x, a.y), fe_const<F>(EdConst::D));
    }
    if (!ok) {
        a.x = fe_zero<F>();
        a.y = fe_one<F>();
        a.t = fe_zero<F>();
        atomicOr(err, 1u);
    }
    ed_aff_store<F>(pts, i, a);
    inf[i] = ok ? 0 : 1;
}

// Scalars must be below the group order r' (canonical Fr); bit 1 of *err
// otherwise.
__global__ void __launch_bounds__(256)
k_bw_check_scalars(const uint32_t* __restrict__ sc, uint32_t n, uint32_t* __restrict__ err) {
    // r' = 0x1cfb69d4...2876e7e1 (banderwagon/cpp/fr.cpp:11-16), little-endian words
    constexpr uint32_t kOrder[8] = {0x2876E7E1u, 0x74FD06B5u, 0x74190471u, 0xFF8F8700u,
                                    0x02687600u, 0x0CCE7602u, 0xCA675F52u, 0x1CFB69D4u};
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint4* s4 = reinterpret_cast<const uint4*>(sc + (size_t)i * 8);
    const uint4 lo = s4[0], hi = s4[1];
    const uint32_t s[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
    uint32_t borrow = 0;  // s - r' borrows exactly when s < r'
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const uint64_t t = (uint64_t)s[k] - kOrder[k] - borrow;
        borrow = (uint32_t)(t >> 63);
    }
    if (!borrow) atomicOr(err, 2u);
}

// The 96-byte output record of an empty MSM: the identity (0, 1) and its
// compressed form (zero bytes).
__global__ void k_bw_identity_record(uint8_t* __restrict__ out) {
    const uint32_t k = threadIdx.x;
    if (k < 96) out[k] = k == 63 ? 1 : 0;
}

void bw_identity_record(uint8_t* out, hipStream_t s) {
    hipLaunchKernelGGL(k_bw_identity_record, dim3(1), dim3(128), 0, s, out);
}

void bw_points_from_wire(const uint8_t* wire, uint32_t* pts, uint8_t* inf, uint32_t n, uint32_t* err,
                         hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_bw_points_from_wire, dim3(ceil_div(n, 128)), dim3(128), 0, s, wire, pts, inf, n, err);
}

void bw_check_scalars(const uint32_t* scalars, uint32_t n, uint32_t* err, hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_bw_check_scalars, dim3(ceil_div(n, 256)), dim3(256), 0, s, scalars, n, err);
}

void launch_accumulate_bw(unsigned grid, hipStream_t s, const uint32_t* entries, const uint32_t* offsets,
                          uint32_t nb, uint32_t chunk, const uint32_t* pts, uint32_t* bsum, uint32_t* carry) {
    hipLaunchKernelGGL(k_accumulate<BwF>, dim3(grid), dim3(64), 0, s, entries, offsets, nb, chunk, pts, bsum, carry);
}

void launch_fold_bw(hipStream_t s, const uint32_t* offsets, uint32_t nb, uint32_t chunk, uint32_t nchunks,
                    uint32_t levels, uint32_t* list, uint32_t* runs, uint32_t* ctl, uint32_t* carry) {
    device_fill_u32(ctl, 0, 3, s);
    uint4* l4 = reinterpret_cast<uint4*>(list);
    uint2* r2 = reinterpret_cast<uint2*>(runs);
    hipLaunchKernelGGL(k_carry_runs<BwF>, dim3(ceil_div((size_t)nb * 4, 256)), dim3(256), 0, s, offsets, nb, chunk,
                       r2, ctl, carry);
    hipLaunchKernelGGL(k_list_runs<BwF>, dim3(256), dim3(256), 0, s, r2, l4, ctl);
    // odd grid stride over the group leaders, as for the other curves
    const unsigned grid = nchunks < 2047 ? (nchunks ? nchunks | 1u : 1u) : 2047;
    uint32_t stride = 1;
    for (uint32_t level = 0; level < levels; level += 6, stride *= kFoldFan)
        hipLaunchKernelGGL(k_fold_groups<BwF>, dim3(grid), dim3(4 * kFoldFan), 0, s, l4, ctl, stride, carry);
}

void launch_reduce_seg_bw(unsigned grid, hipStream_t s, const uint32_t* offsets, uint32_t chunk, const uint32_t* bsum,
                          const uint32_t* carry, uint32_t nseg, uint32_t seg, uint32_t* segA, uint32_t* segB) {
    hipLaunchKernelGGL(k_reduce_seg<BwF>, dim3(grid), dim3(256), 0, s, offsets, chunk, bsum, carry, nseg, seg, segA,
                       segB);
}

void launch_reduce_level_bw(unsigned grid, hipStream_t s, const uint32_t* inA, const uint32_t* inB,
                            const RedLevel* tab, uint32_t ntab, uint32_t uni_n, uint32_t uni_parts, uint32_t lm,
                            uint32_t* outA, uint32_t* outB) {
    hipLaunchKernelGGL(k_reduce_level<BwF>, dim3(grid), dim3(kLevelThreads), 0, s, inA, inB, tab, ntab, uni_n,
                       uni_parts, lm, outA, outB);
}

}  // namespace kz
