// Window combine with row-distributed field arithmetic (field_row.h): XYZZ
// points for BLS12-381, BN254 and secp256k1, extended Edwards points for
// Banderwagon.  One wave per MSM; its four 16-lane rows evaluate a level of
// the point formula together, each product spread over a row's lanes.  The
// chain of c (W - 1) doublings is the critical path of a single MSM's tail;
// this cuts the issue time of each of its products from 2 N^2 multiply-adds
// to about 4 N per lane.
//
// Built with the DPP-combine pass off (Makefile): it would fold the row_ror /
// row_newbcast moves into the instructions that use them, and a folded DPP
// subtraction once gave wrong differences in this library's quad code.
#include "field_row.h"
#include "msm_impl.h"
#include "msm_stages.h"

namespace kz {

// Lane j (< N) of every row reads limb j of each coordinate of record idx.
template <class F>
__device__ __forceinline__ XyzzRow<WaveDev, F> row_load(const uint32_t* __restrict__ base, size_t idx) {
    constexpr int K = Rec<F>::kXyzzWords;
    const uint32_t j = threadIdx.x & 15u;
    const uint32_t* rec = base + idx * K;
    XyzzRow<WaveDev, F> r;
    const bool in = j < (uint32_t)F::N;
    r.X.v = in ? rec[j] : 0u;
    r.Y.v = in ? rec[F::N + j] : 0u;
    r.ZZ.v = in ? rec[2 * F::N + j] : 0u;
    r.ZZZ.v = in ? rec[3 * F::N + j] : 0u;
    return r;
}

// Horner over the window sums, MSW first: r = 2^c r + S_w.  Block b (one
// wave) combines job b; jobs = null: a single MSM described by `one`.
template <class C>
__global__ void __launch_bounds__(64) k_combine_row(const uint32_t* __restrict__ winres,
                                                    const MsmJob* __restrict__ jobs, MsmJob one, uint32_t njobs,
                                                    uint8_t* __restrict__ out) {
    using F = typename C::Fp;
    using W = WaveDev;
    const uint32_t job = blockIdx.x;
    if (job >= njobs) return;
    const MsmJob jd = jobs ? jobs[job] : one;
    const RowConsts<W, F> K = row_consts<W, F>();
    const int c = (int)jd.c;
    XyzzRow<W, F> r = row_load<F>(winres, jd.win_base + jd.nwin - 1);
#pragma unroll 1
    for (int w = (int)jd.nwin - 2; w >= 0; --w) {
        if (!row_pt_is_inf<W, F>(r)) {
#pragma unroll 1
            for (int i = 0; i < c; ++i) r = row_dbl<W, F>(r, K);
        }
        r = row_add<W, F>(r, row_load<F>(winres, jd.win_base + (uint32_t)w), K);
    }
    const Xyzz<F> s = row_to_xyzz<W, F>(r);
    if (threadIdx.x == 0) write_output<C>(xyzz_canon_acc<F>(s), out + (size_t)job * 3 * C::FB);
}

// Banderwagon: extended Edwards records (X, Y, T, Z of N words each), the
// complete formulas of ed_dbl_quad_lazy / ed_add_quad_lazy by rows.
template <class F>
__device__ __forceinline__ EdRow<WaveDev, F> row_load_ed(const uint32_t* __restrict__ base, size_t idx) {
    const uint32_t j = threadIdx.x & 15u;
    const uint32_t* rec = base + idx * EdRec::kPtWords;
    EdRow<WaveDev, F> r;
    const bool in = j < (uint32_t)F::N;
    r.X.v = in ? rec[j] : 0u;
    r.Y.v = in ? rec[F::N + j] : 0u;
    r.T.v = in ? rec[2 * F::N + j] : 0u;
    r.Z.v = in ? rec[3 * F::N + j] : 0u;
    return r;
}

__global__ void __launch_bounds__(64) k_combine_row_ed(const uint32_t* __restrict__ winres,
                                                       const MsmJob* __restrict__ jobs, MsmJob one, uint32_t njobs,
                                                       uint8_t* __restrict__ out) {
    using C = CurveBanderwagon;
    using F = C::Fp;
    using W = WaveDev;
    const uint32_t job = blockIdx.x;
    if (job >= njobs) return;
    const MsmJob jd = jobs ? jobs[job] : one;
    const RowConsts<W, F> K = row_consts<W, F>();
    const FeRow<W, F> d = row_from_fe<W, F>(fe_const<F>(EdConst::D));
    const int c = (int)jd.c;
    EdRow<W, F> r = row_load_ed<F>(winres, jd.win_base + jd.nwin - 1);
#pragma unroll 1
    for (int w = (int)jd.nwin - 2; w >= 0; --w) {
        if (!row_ed_is_identity<W, F>(r)) {
#pragma unroll 1
            for (int i = 0; i < c; ++i) r = row_ed_dbl<W, F>(r, K);
        }
        r = row_ed_add<W, F>(r, row_load_ed<F>(winres, jd.win_base + (uint32_t)w), d, K);
    }
    const EdExt<F> s = row_to_ed<W, F>(r);
    if (threadIdx.x == 0) write_output<C>(ed_canon<F>(s), out + (size_t)job * 3 * C::FB);
}

void launch_combine_bw(hipStream_t s, const uint32_t* winres, const MsmJob* jobs, const MsmJob& one, uint32_t njobs,
                       uint8_t* out) {
    hipLaunchKernelGGL(k_combine_row_ed, dim3(njobs), dim3(64), 0, s, winres, jobs, one, njobs, out);
}

void launch_combine_bls(hipStream_t s, const uint32_t* winres, const MsmJob* jobs, const MsmJob& one,
                        uint32_t njobs, uint8_t* out) {
    hipLaunchKernelGGL(k_combine_row<CurveBLS12381>, dim3(njobs), dim3(64), 0, s, winres, jobs, one, njobs, out);
}

void launch_combine_bn(hipStream_t s, const uint32_t* winres, const MsmJob* jobs, const MsmJob& one, uint32_t njobs,
                       uint8_t* out) {
    hipLaunchKernelGGL(k_combine_row<CurveBN254>, dim3(njobs), dim3(64), 0, s, winres, jobs, one, njobs, out);
}

void launch_combine_secp(hipStream_t s, const uint32_t* winres, const MsmJob* jobs, const MsmJob& one,
                         uint32_t njobs, uint8_t* out) {
    hipLaunchKernelGGL(k_combine_row<CurveSecp256k1>, dim3(njobs), dim3(64), 0, s, winres, jobs, one, njobs, out);
}

}  // namespace kz
