// Fixed-base batch scalar multiplication (kernels in gen_impl.h) and its C ABI.
#include <cstring>

#include "engine.h"
#include "gen_impl.h"
#include "kinet/crypto/hip_runtime.h"

namespace kz {

// Table of d * 2^(8j) * G, built on first use per curve and kept resident.
static const uint32_t* gen_table(Engine& e, int curve, hipStream_t s) {
    DevBuf& t = e.gen_table[curve];
    if (t.ptr) return t.as<uint32_t>();
    const size_t entries = (size_t)kGenWindows * kGenDigits;
    const size_t xw = (size_t)xyzz_words(curve), aw = (size_t)affine_words(curve);
    DevBuf scratch;
    scratch.ensure((kGenWindows + entries) * xw * 4);
    t.ensure(entries * aw * 4);
    if (curve == 3) gen_build_table<CurveBanderwagon>(scratch.as<uint32_t>(), t.as<uint32_t>(), s);
    else if (curve == 2) gen_build_table<CurveBLS12381>(scratch.as<uint32_t>(), t.as<uint32_t>(), s);
    else if (curve == 1) gen_build_table<CurveBN254>(scratch.as<uint32_t>(), t.as<uint32_t>(), s);
    else gen_build_table<CurveSecp256k1>(scratch.as<uint32_t>(), t.as<uint32_t>(), s);
    KZ_HIP(hipGetLastError());
    KZ_HIP(hipStreamSynchronize(s));  // scratch is freed on return
    return t.as<uint32_t>();
}

void fixed_base_mul(int curve, const uint32_t* d_scalars, size_t n, uint8_t* d_out_wire, hipStream_t s) {
    Engine& e = engine();
    std::lock_guard<std::mutex> g(e.mu);
    const uint32_t* table = gen_table(e, curve, s);
    if (curve == 3) gen_mul<CurveBanderwagon>(table, d_scalars, (uint32_t)n, d_out_wire, s);
    else if (curve == 2) gen_mul<CurveBLS12381>(table, d_scalars, (uint32_t)n, d_out_wire, s);
    else if (curve == 1) gen_mul<CurveBN254>(table, d_scalars, (uint32_t)n, d_out_wire, s);
    else gen_mul<CurveSecp256k1>(table, d_scalars, (uint32_t)n, d_out_wire, s);
    KZ_HIP(hipGetLastError());
}

}  // namespace kz

extern "C" int kinet_hip_g1_fixed_base_mul(uint32_t curve, const void* d_scalars, size_t n, void* d_out_wire,
                                           void* stream) {
    if (curve > 3) return -3;
    if (n > 0 && (!d_scalars || !d_out_wire)) return -1;
    if (n >= (size_t)1 << 31) return -2;
    try {
        kz::fixed_base_mul((int)curve, static_cast<const uint32_t*>(d_scalars), n, static_cast<uint8_t*>(d_out_wire),
                           static_cast<hipStream_t>(stream));
        return 0;
    } catch (const std::exception& ex) {
        kz::set_last_error(ex.what());
        return -3;
    }
}
