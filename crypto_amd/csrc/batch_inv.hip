// Montgomery batch inversion over the three base fields on the device, with
// the contract of gpukit/include/kinet/gpukit/batch_inversion.h:1-15:
// inv[i] = a[i]^-1 mod p for little-endian elements of 32 bytes (secp256k1,
// BN254) or 48 bytes (BLS12-381), GPUKIT_ERR_BAD_SIZE when an element is
// zero.  An element is read as its residue mod p, so a nonzero multiple of p
// is a zero element too.  The reference's CUDA form is a stub (gpukit/gpu/
// cuda/batch_inversion.cu:6-14) and its CPU form
// (gpukit/cpp/cpu_reference/batch_inversion.cpp:166-194, 312-338) does not
// return inverses: its reduction u512_mod_p_generic / u768_mod_p_generic
// (:110-141, :263-289) subtracts p 2^k whenever a bit is set, with no
// comparison, so the results follow the header's arithmetic, not those bytes.
//
// Device schedule.  T threads each run Montgomery's trick over the elements
// t, t + T, t + 2T, ... (neighbouring threads read neighbouring records):
// prefix products a_t a_(t+T) ... (one product per element, stored in the
// output slots as scratch), one inversion of the thread's total (divsteps),
// then a backward walk: out_i = inv * prefix_(i-T), inv <- inv * a_i (two
// products per element).  T is about n / 16, so one inversion is shared by
// ~16 elements; three products per element remain.
#include <algorithm>

#include "engine.h"
#include "field.h"
#include "kinet/crypto/hip_runtime.h"
#include "kinet/gpukit/batch_inversion.h"
#include "kinet/gpukit/gpukit.h"
#include "scan.h"

namespace kz {
namespace {

constexpr uint32_t kInvZero = 1u;  // an element is zero mod p

template <class F, int NW>
KZ_DEV Fe<F> load_elem(const uint32_t* p) {
    uint32_t w[NW];
#pragma unroll
    for (int k = 0; k < NW; ++k) w[k] = p[k];
    return fe_from_words<F>(w, NW);
}

template <class F, int NW>
KZ_DEV void store_elem(uint32_t* p, const Fe<F>& a) {
    uint32_t w[NW];
    fe_to_words<F>(a, w, NW);
#pragma unroll
    for (int k = 0; k < NW; ++k) p[k] = w[k];
}

template <class F, int NW>
__global__ void __launch_bounds__(256)
k_batch_inv(const uint32_t* __restrict__ in, uint32_t* __restrict__ out, size_t n, size_t T,
            uint32_t* __restrict__ status) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T || t >= n) return;
    Fe<F> acc = fe_one<F>();
    uint32_t flag = 0;
    size_t last = t;
#pragma unroll 1
    for (size_t i = t; i < n; i += T) {
        const Fe<F> v = fe_to_mont<F>(load_elem<F, NW>(in + i * NW));  // residue, Montgomery form
        if (fe_is_zero<F>(v)) flag = kInvZero;
        acc = fe_mul<F>(acc, v);
        store_elem<F, NW>(out + i * NW, acc);  // prefix product (scratch)
        last = i;
    }
    if (flag) {  // a zero element: the outputs are refused
        atomicOr(status, flag);
        return;
    }
    Fe<F> inv = fe_inv<F>(acc);
#pragma unroll 1
    for (size_t i = last;; i -= T) {
        const Fe<F> r = (i >= T) ? fe_mul<F>(inv, load_elem<F, NW>(out + (i - T) * NW)) : inv;
        if (i >= T) inv = fe_mul<F>(inv, fe_to_mont<F>(load_elem<F, NW>(in + i * NW)));
        store_elem<F, NW>(out + i * NW, fe_from_mont<F>(r));
        if (i < T) break;
    }
}

int words_of(uint32_t curve) { return curve == 2 ? 12 : 8; }

// Threads: about 16 elements each, at most 2^17 threads.
size_t threads_for(size_t n) { return std::min<size_t>(std::max<size_t>(1, (n + 15) / 16), (size_t)1 << 17); }

void launch(uint32_t curve, const uint32_t* d_in, uint32_t* d_out, size_t n, uint32_t* d_status, hipStream_t s) {
    device_fill_u32(d_status, 0, 1, s);
    const size_t T = threads_for(n);
    const dim3 grid((unsigned)ceil_div(T, 256)), block(256);
    if (curve == 0) hipLaunchKernelGGL((k_batch_inv<SecpFp, 8>), grid, block, 0, s, d_in, d_out, n, T, d_status);
    else if (curve == 1) hipLaunchKernelGGL((k_batch_inv<BnFp, 8>), grid, block, 0, s, d_in, d_out, n, T, d_status);
    else hipLaunchKernelGGL((k_batch_inv<BlsFp, 12>), grid, block, 0, s, d_in, d_out, n, T, d_status);
    KZ_HIP(hipGetLastError());
}

int batch_inv_host(uint32_t curve, const uint8_t* in, uint8_t* out, size_t n) {
    if (n == 0) return GPUKIT_OK;
    if (!in || !out) return GPUKIT_ERR_NULL_ARG;
    if (n > ((size_t)1 << 32)) return GPUKIT_ERR_BAD_SIZE;
    try {
        Engine& e = engine();
        std::lock_guard<std::mutex> g(e.mu);
        hipStream_t s = e.stream;
        const size_t bytes = n * 4 * (size_t)words_of(curve);
        uint32_t* d_in = static_cast<uint32_t*>(e.h2d_a.ensure(bytes));
        uint32_t* d_out = static_cast<uint32_t*>(e.h2d_b.ensure(bytes));
        uint32_t* d_status = static_cast<uint32_t*>(e.ntt_tw.ensure(sizeof(uint32_t)));
        KZ_HIP(hipMemcpyAsync(d_in, in, bytes, hipMemcpyHostToDevice, s));
        launch(curve, d_in, d_out, n, d_status, s);
        uint32_t status = 0;
        KZ_HIP(hipMemcpyAsync(&status, d_status, sizeof status, hipMemcpyDeviceToHost, s));
        KZ_HIP(hipStreamSynchronize(s));
        if (status & kInvZero) return GPUKIT_ERR_BAD_SIZE;  // out untouched
        KZ_HIP(hipMemcpyAsync(out, d_out, bytes, hipMemcpyDeviceToHost, s));
        KZ_HIP(hipStreamSynchronize(s));
        return GPUKIT_OK;
    } catch (const std::exception& ex) {
        set_last_error(ex.what());
        return GPUKIT_ERR_BACKEND;
    }
}

}  // namespace
}  // namespace kz

using namespace kz;

extern "C" {

int gpukit_batch_inv_secp256k1_fp_hip(const uint8_t* in, uint8_t* out, size_t n) {
    return batch_inv_host(0, in, out, n);
}

int gpukit_batch_inv_bn254_fp_hip(const uint8_t* in, uint8_t* out, size_t n) { return batch_inv_host(1, in, out, n); }

int gpukit_batch_inv_bls12_381_fp_hip(const uint8_t* in, uint8_t* out, size_t n) {
    return batch_inv_host(2, in, out, n);
}

int kinet_hip_batch_inv_fp(uint32_t curve, const void* d_in, void* d_out, size_t n, uint32_t* d_status,
                           void* stream) {
    if (curve > 2) return GPUKIT_ERR_BACKEND;
    if (n == 0) return GPUKIT_OK;
    if (!d_in || !d_out || !d_status || !stream) return GPUKIT_ERR_NULL_ARG;
    const size_t bytes = n * 4 * (size_t)words_of(curve);
    const uintptr_t a = reinterpret_cast<uintptr_t>(d_in), b = reinterpret_cast<uintptr_t>(d_out);
    if (n > ((size_t)1 << 32) || ((a | b) & 3u) != 0 || (a < b + bytes && b < a + bytes)) return GPUKIT_ERR_BAD_SIZE;
    try {
        launch(curve, static_cast<const uint32_t*>(d_in), static_cast<uint32_t*>(d_out), n, d_status,
               static_cast<hipStream_t>(stream));
        return GPUKIT_OK;
    } catch (const std::exception& ex) {
        set_last_error(ex.what());
        return GPUKIT_ERR_BACKEND;
    }
}

}  // extern "C"
