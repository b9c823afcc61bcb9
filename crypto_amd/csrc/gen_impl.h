// Fixed-base batch scalar multiplication k_i * G on the device (synthetic MSM
// points, SRS points [tau^i] G).  Included by one translation unit per curve.
//
// Table T[j][d] = d * 2^(8j) * G for 32 byte positions j and d = 1..255
// (8160 affine points, ~0.9 MB for BLS12-381), built once per curve:
//   k_gen_bases    one thread: B_j = 2^(8j) G by 8 doublings per step
//   k_gen_table    one thread per (j, d): d * B_j by double-and-add
//   k_gen_norm     one thread per entry: affine normalisation
// then k_gen_mul adds T[j][byte_j(k)] over the 32 bytes of each scalar (any
// 256-bit k: the bytes reconstruct k exactly) and writes the affine point in
// the wire format (BE x || y, identity = zero bytes).  The point operations
// are the curve's (point_ops.h): XYZZ, or extended Edwards for Banderwagon.
#pragma once

#include "point_ops.h"

namespace kz {

constexpr int kGenWindows = 32;
constexpr int kGenDigits = 256;

template <class C>
__device__ typename PointOps<typename C::Fp>::Aff generator_affine() {
    using F = typename C::Fp;
    // BLS12-381 G1 generator: bls/cpp/bls12_381_g1.hpp:204-224; BN254: (1, 2)
    // (gpukit/test/multi_pippenger_test.cpp:207-211); secp256k1: SEC 2's G
    // (secp256k1/cpp/field.hpp:88-91).  Canonical, big-endian.
    constexpr uint8_t bls_x[48] = {
        0x17, 0xF1, 0xD3, 0xA7, 0x31, 0x97, 0xD7, 0x94, 0x26, 0x95, 0x63, 0x8C, 0x4F, 0xA9, 0xAC, 0x0F,
        0xC3, 0x68, 0x8C, 0x4F, 0x97, 0x74, 0xB9, 0x05, 0xA1, 0x4E, 0x3A, 0x3F, 0x17, 0x1B, 0xAC, 0x58,
        0x6C, 0x55, 0xE8, 0x3F, 0xF9, 0x7A, 0x1A, 0xEF, 0xFB, 0x3A, 0xF0, 0x0A, 0xDB, 0x22, 0xC6, 0xBB};
    constexpr uint8_t bls_y[48] = {
        0x08, 0xB3, 0xF4, 0x81, 0xE3, 0xAA, 0xA0, 0xF1, 0xA0, 0x9E, 0x30, 0xED, 0x74, 0x1D, 0x8A, 0xE4,
        0xFC, 0xF5, 0xE0, 0x95, 0xD5, 0xD0, 0x0A, 0xF6, 0x00, 0xDB, 0x18, 0xCB, 0x2C, 0x04, 0xB3, 0xED,
        0xD0, 0x3C, 0xC7, 0x44, 0xA2, 0x88, 0x8A, 0xE4, 0x0C, 0xAA, 0x23, 0x29, 0x46, 0xC5, 0xE7, 0xE1};
    constexpr uint8_t secp_x[32] = {
        0x79, 0xBE, 0x66, 0x7E, 0xF9, 0xDC, 0xBB, 0xAC, 0x55, 0xA0, 0x62, 0x95, 0xCE, 0x87, 0x0B, 0x07,
        0x02, 0x9B, 0xFC, 0xDB, 0x2D, 0xCE, 0x28, 0xD9, 0x59, 0xF2, 0x81, 0x5B, 0x16, 0xF8, 0x17, 0x98};
    constexpr uint8_t secp_y[32] = {
        0x48, 0x3A, 0xDA, 0x77, 0x26, 0xA3, 0xC4, 0x65, 0x5D, 0xA4, 0xFB, 0xFC, 0x0E, 0x11, 0x08, 0xA8,
        0xFD, 0x17, 0xB4, 0x48, 0xA6, 0x85, 0x54, 0x19, 0x9C, 0x47, 0xD0, 0x8F, 0xFB, 0x10, 0xD4, 0xB8};
    // Banderwagon: the Bandersnatch base point of banderwagon/cpp/element.cpp:48-75
    constexpr uint8_t bw_x[32] = {
        0x29, 0xC1, 0x32, 0xCC, 0x2C, 0x0B, 0x34, 0xC5, 0x74, 0x37, 0x11, 0x77, 0x7B, 0xBE, 0x42, 0xF3,
        0x2B, 0x79, 0xC0, 0x22, 0xAD, 0x99, 0x84, 0x65, 0xE1, 0xE7, 0x18, 0x66, 0xA2, 0x52, 0xAE, 0x18};
    constexpr uint8_t bw_y[32] = {
        0x2A, 0x6C, 0x66, 0x9E, 0xDA, 0x12, 0x3E, 0x0F, 0x15, 0x7D, 0x8B, 0x50, 0xBA, 0xDC, 0xD5, 0x86,
        0x35, 0x8C, 0xAD, 0x81, 0xEE, 0xE4, 0x64, 0x60, 0x5E, 0x31, 0x67, 0xB6, 0xCC, 0x97, 0x41, 0x66};
    if constexpr (C::ID == 3) {
        EdAff<F> g;
        g.x = fe_to_mont<F>(fe_from_be_bytes<F>(bw_x, 32));
        g.y = fe_to_mont<F>(fe_from_be_bytes<F>(bw_y, 32));
        g.t = fe_mul<F>(fe_mul<F>(g.x, g.y), fe_const<F>(EdConst::D));
        return g;
    } else {
        Affine<F> g;
        if constexpr (C::ID == 2) {
            g.x = fe_to_mont<F>(fe_from_be_bytes<F>(bls_x, 48));
            g.y = fe_to_mont<F>(fe_from_be_bytes<F>(bls_y, 48));
        } else if constexpr (C::ID == 0) {
            g.x = fe_to_mont<F>(fe_from_be_bytes<F>(secp_x, 32));
            g.y = fe_to_mont<F>(fe_from_be_bytes<F>(secp_y, 32));
        } else {
            Fe<F> x = fe_zero<F>(), y = fe_zero<F>();
            x.v[0] = 1;
            y.v[0] = 2;
            g.x = fe_to_mont<F>(x);
            g.y = fe_to_mont<F>(y);
        }
        return g;
    }
}

template <class C>
__global__ void k_gen_bases(uint32_t* __restrict__ bases) {
    using P = PointOps<typename C::Fp>;
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    typename P::Pt b = P::from_aff(generator_affine<C>());
#pragma unroll 1
    for (int j = 0; j < kGenWindows; ++j) {
        P::store(bases, j, b);
#pragma unroll 1
        for (int i = 0; i < 8; ++i) b = P::template dbl<MulCall>(b);
    }
}

template <class C>
__global__ void __launch_bounds__(64) k_gen_table(const uint32_t* __restrict__ bases, uint32_t* __restrict__ table) {
    using P = PointOps<typename C::Fp>;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;  // t = j * 256 + d
    if (t >= kGenWindows * kGenDigits) return;
    const uint32_t j = t >> 8, d = t & 255;
    const typename P::Pt b = P::load(bases, j);
    typename P::Pt acc = P::inf();
#pragma unroll 1
    for (int bit = 7; bit >= 0; --bit) {
        acc = P::template dbl<MulCall>(acc);
        if ((d >> bit) & 1u) acc = P::template add<MulCall>(acc, b);
    }
    P::store(table, t, acc);
}

template <class C>
__global__ void __launch_bounds__(64) k_gen_norm(const uint32_t* __restrict__ table, uint32_t* __restrict__ aff) {
    using P = PointOps<typename C::Fp>;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= kGenWindows * kGenDigits) return;
    typename P::Aff a;
    bool inf;
    P::to_affine(P::load(table, t), a, inf);  // d = 0 entries are the identity, never read
    P::store_aff(aff, t, a);
}

template <class C>
__global__ void __launch_bounds__(64)
k_gen_mul(const uint32_t* __restrict__ aff_table, const uint32_t* __restrict__ scalars, uint32_t n,
          uint8_t* __restrict__ out_wire) {
    using F = typename C::Fp;
    using P = PointOps<F>;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t* s = scalars + (size_t)i * 8;
    typename P::Pt acc = P::inf();
#pragma unroll 1
    for (int j = 0; j < kGenWindows; ++j) {
        const uint32_t d = (s[j >> 2] >> (8 * (j & 3))) & 255u;  // cached global read, no register indexing
        if (d) P::template madd<MulCall>(acc, P::load_aff(aff_table, (uint32_t)j * kGenDigits + d));
    }
    typename P::Aff a;
    bool inf;
    P::to_affine(acc, a, inf);
    uint8_t* w = out_wire + (size_t)i * 2 * C::FB;
    if (inf) {
        for (int k = 0; k < 2 * C::FB; ++k) w[k] = 0;
        return;
    }
    fe_to_be_bytes<F>(fe_from_mont<F>(a.x), w, C::FB);
    fe_to_be_bytes<F>(fe_from_mont<F>(a.y), w + C::FB, C::FB);
}

// Builds the table into `aff_table` (kGenWindows * kGenDigits affine records),
// using `scratch` (kGenWindows + kGenWindows * kGenDigits XYZZ records).
template <class C>
void gen_build_table(uint32_t* scratch, uint32_t* aff_table, hipStream_t s) {
    uint32_t* bases = scratch;
    uint32_t* table = scratch + (size_t)kGenWindows * PointOps<typename C::Fp>::kPtWords;
    const unsigned nt = kGenWindows * kGenDigits;
    hipLaunchKernelGGL(k_gen_bases<C>, dim3(1), dim3(64), 0, s, bases);
    hipLaunchKernelGGL(k_gen_table<C>, dim3(nt / 64), dim3(64), 0, s, bases, table);
    hipLaunchKernelGGL(k_gen_norm<C>, dim3(nt / 64), dim3(64), 0, s, table, aff_table);
}

template <class C>
void gen_mul(const uint32_t* aff_table, const uint32_t* d_scalars, uint32_t n, uint8_t* d_out, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_gen_mul<C>, dim3((n + 63) / 64), dim3(64), 0, s, aff_table, d_scalars, n, d_out);
}

}  // namespace kz
