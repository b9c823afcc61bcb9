// BN254 unit operations on the device in the reference's batch layout:
// kinet_bn254_hip_{available, g1_add, g1_mul, fp_mul}, the HIP counterparts of
// kinet_bn254_cuda_{available, g1_add, g1_mul, fp_mul}
// (bn254/gpu/cuda/bn254_driver_cuda.h:13-28).  They expose the field and
// point arithmetic the MSM runs on, one operation per element, so it can be
// checked in isolation against the reference's CPU code.
//
// Layout (as the reference driver): an affine point is 9 x u64 LE, x[4] ||
// y[4] || inf (inf != 0 = the identity), coordinates in the reference's
// Montgomery form x R mod p with R = 2^256 (bn254/cpp/bn254_fp.hpp); a scalar
// is 4 x u64 LE, any 256-bit value; an Fp element is 4 x u64 LE, Montgomery.
//
// The device works in radix-2^29 limbs with R' = 2^261.  Conversions: a
// reference value A = x 2^256 becomes x 2^261 = mont'(A, 2^266 mod p), and
// back, x 2^256 = mont'(x 2^261, 2^256 mod p).  Both constants are made on
// the device: 2^266 mod p = to_mont'(32), 2^256 mod p = mont'(to_mont'(2^128),
// 2^128).
#include <cstring>

#include "curve.h"
#include "engine.h"
#include "kinet/crypto/bn254.h"
#include "kinet/crypto/hip_runtime.h"

namespace kz {
namespace {

using F = BnFp;

struct RefForm {
    Fe<F> k_in;   // 2^266 mod p
    Fe<F> k_out;  // 2^256 mod p
};

KZ_DEV Fe<F> fe_small(uint32_t v) {
    Fe<F> a = fe_zero<F>();
    a.v[0] = v;
    return a;
}

KZ_DEV RefForm ref_form() {
    RefForm r;
    r.k_in = fe_to_mont<F>(fe_small(32));
    uint32_t w[8] = {0, 0, 0, 0, 1, 0, 0, 0};  // 2^128
    const Fe<F> t128 = fe_from_words<F>(w, 8);
    r.k_out = fe_mul<F>(fe_to_mont<F>(t128), t128);
    return r;
}

// 4 x u64 LE (reference Montgomery form) -> device Montgomery form
KZ_DEV Fe<F> load_ref(const uint64_t* p, const RefForm& rf) {
    uint32_t w[8];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        w[2 * i] = (uint32_t)p[i];
        w[2 * i + 1] = (uint32_t)(p[i] >> 32);
    }
    return fe_mul<F>(fe_from_words<F>(w, 8), rf.k_in);
}

KZ_DEV void store_ref(uint64_t* p, const Fe<F>& a, const RefForm& rf) {
    uint32_t w[8];
    fe_to_words<F>(fe_mul<F>(a, rf.k_out), w, 8);
#pragma unroll
    for (int i = 0; i < 4; ++i) p[i] = (uint64_t)w[2 * i] | ((uint64_t)w[2 * i + 1] << 32);
}

KZ_DEV void store_point(uint64_t* p, const Xyzz<F>& q, const RefForm& rf) {
    Affine<F> a;
    bool inf;
    xyzz_to_affine<F>(q, a, inf);
    if (inf) {
#pragma unroll
        for (int i = 0; i < 8; ++i) p[i] = 0;
        p[8] = 1;
        return;
    }
    store_ref(p, a.x, rf);
    store_ref(p + 4, a.y, rf);
    p[8] = 0;
}

__global__ void __launch_bounds__(64) k_bn_fp_mul(const uint64_t* __restrict__ a, const uint64_t* __restrict__ b,
                                                  uint64_t* __restrict__ out, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const RefForm rf = ref_form();
    store_ref(out + 4 * (size_t)i, fe_mul<F>(load_ref(a + 4 * (size_t)i, rf), load_ref(b + 4 * (size_t)i, rf)), rf);
}

__global__ void __launch_bounds__(64) k_bn_g1_add(const uint64_t* __restrict__ a, const uint64_t* __restrict__ b,
                                                  uint64_t* __restrict__ out, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const RefForm rf = ref_form();
    const uint64_t* pa = a + 9 * (size_t)i;
    const uint64_t* pb = b + 9 * (size_t)i;
    Xyzz<F> acc = xyzz_inf<F>();
    if (pa[8] == 0) {
        Affine<F> A{load_ref(pa, rf), load_ref(pa + 4, rf)};
        acc = xyzz_from_affine<F>(A);
    }
    if (pb[8] == 0) {
        Affine<F> B{load_ref(pb, rf), load_ref(pb + 4, rf)};
        xyzz_madd<F, MulCall>(acc, B);
    }
    store_point(out + 9 * (size_t)i, acc, rf);
}

// k P by double-and-add from the top bit (variable time: a test and
// inspection entry, not for secret scalars).
__global__ void __launch_bounds__(64) k_bn_g1_mul(const uint64_t* __restrict__ pts, const uint64_t* __restrict__ sc,
                                                  uint64_t* __restrict__ out, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const RefForm rf = ref_form();
    const uint64_t* pp = pts + 9 * (size_t)i;
    Xyzz<F> acc = xyzz_inf<F>();
    if (pp[8] == 0) {
        const Affine<F> P{load_ref(pp, rf), load_ref(pp + 4, rf)};
        const uint64_t* k = sc + 4 * (size_t)i;
        for (int bit = 255; bit >= 0; --bit) {
            acc = xyzz_dbl<F, MulCall>(acc);
            if ((k[bit >> 6] >> (bit & 63)) & 1) xyzz_madd<F, MulCall>(acc, P);
        }
    }
    store_point(out + 9 * (size_t)i, acc, rf);
}

// Host buffers in, host buffers out: stage through the engine's buffers,
// launch one thread per element, copy back, synchronise.
template <class Launch>
int run_batch(const void* a, size_t a_bytes, const void* b, size_t b_bytes, void* out, size_t out_bytes,
              Launch&& launch) {
    try {
        Engine& e = engine();
        std::lock_guard<std::mutex> g(e.mu);
        hipStream_t s = e.stream;
        void* da = e.h2d_a.ensure(a_bytes);
        void* db = e.h2d_b.ensure(b_bytes);
        void* dout = e.out.ensure(out_bytes);
        KZ_HIP(hipMemcpyAsync(da, a, a_bytes, hipMemcpyHostToDevice, s));
        KZ_HIP(hipMemcpyAsync(db, b, b_bytes, hipMemcpyHostToDevice, s));
        launch(static_cast<const uint64_t*>(da), static_cast<const uint64_t*>(db), static_cast<uint64_t*>(dout), s);
        KZ_HIP(hipGetLastError());
        KZ_HIP(hipMemcpyAsync(out, dout, out_bytes, hipMemcpyDeviceToHost, s));
        KZ_HIP(hipStreamSynchronize(s));
        return 0;
    } catch (const std::exception& ex) {
        set_last_error(ex.what());
        return -1;
    }
}

}  // namespace
}  // namespace kz

using namespace kz;

extern "C" {

int kinet_bn254_hip_available(void) { return kinet_hip_available(); }

int kinet_bn254_hip_fp_mul(const void* a, const void* b, void* out, unsigned n) {
    if (n == 0) return 0;
    if (!a || !b || !out) return -1;
    const size_t bytes = (size_t)32 * n;
    return run_batch(a, bytes, b, bytes, out, bytes, [n](const uint64_t* da, const uint64_t* db, uint64_t* dout,
                                                        hipStream_t s) {
        hipLaunchKernelGGL(k_bn_fp_mul, dim3(ceil_div(n, 64)), dim3(64), 0, s, da, db, dout, (uint32_t)n);
    });
}

int kinet_bn254_hip_g1_add(const void* a, const void* b, void* out, unsigned n) {
    if (n == 0) return 0;
    if (!a || !b || !out) return -1;
    const size_t bytes = (size_t)72 * n;
    return run_batch(a, bytes, b, bytes, out, bytes, [n](const uint64_t* da, const uint64_t* db, uint64_t* dout,
                                                        hipStream_t s) {
        hipLaunchKernelGGL(k_bn_g1_add, dim3(ceil_div(n, 64)), dim3(64), 0, s, da, db, dout, (uint32_t)n);
    });
}

int kinet_bn254_hip_g1_mul(const void* points, const void* scalars, void* out, unsigned n) {
    if (n == 0) return 0;
    if (!points || !scalars || !out) return -1;
    return run_batch(points, (size_t)72 * n, scalars, (size_t)32 * n, out, (size_t)72 * n,
                     [n](const uint64_t* dp, const uint64_t* ds, uint64_t* dout, hipStream_t s) {
                         hipLaunchKernelGGL(k_bn_g1_mul, dim3(ceil_div(n, 64)), dim3(64), 0, s, dp, ds, dout,
                                            (uint32_t)n);
                     });
}

}  // extern "C"
