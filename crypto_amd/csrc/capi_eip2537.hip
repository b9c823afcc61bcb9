// EIP-2537 G1 MSM entry (include/kinet/crypto/bls.h).
//
// Same checks and result as the reference (bls/cpp/bls.cpp:194-245): each
// coordinate a 64-byte big-endian integer below p, each point on the curve and
// in G1, infinity (0, 0) skipped, full 256-bit big-endian scalars.  The
// encodings are checked on the host; the points then go to the device in the
// gpukit wire layout (48-byte big-endian x || y, all zero = infinity), are
// checked there by the G1 kernel, and the MSM runs on the G1 path (GLV split).
#include <cstring>
#include <vector>

#include "engine.h"
#include "glv.h"
#include "kinet/crypto/bls.h"
#include "scan.h"

namespace kz {
namespace {

constexpr size_t kEntry = 64 + 64 + 32;
// BLS12-381 p, 48 bytes big-endian
constexpr uint8_t kPBe[48] = {0x1a, 0x01, 0x11, 0xea, 0x39, 0x7f, 0xe6, 0x9a, 0x4b, 0x1b, 0xa7, 0xb6,
                              0x43, 0x4b, 0xac, 0xd7, 0x64, 0x77, 0x4b, 0x84, 0xf3, 0x85, 0x12, 0xbf,
                              0x67, 0x30, 0xd2, 0xa0, 0xf6, 0xb0, 0xf6, 0x24, 0x1e, 0xab, 0xff, 0xfe,
                              0xb1, 0x53, 0xff, 0xff, 0xb9, 0xfe, 0xff, 0xff, 0xff, 0xff, 0xaa, 0xab};

// a 64-byte big-endian integer below p
bool fp_encoding_ok(const uint8_t* c) {
    for (int k = 0; k < 16; ++k)
        if (c[k]) return false;
    return std::memcmp(c + 16, kPBe, 48) < 0;
}

}  // namespace
}  // namespace kz

using namespace kz;

extern "C" int bls12_381_g1_msm(uint8_t rx[64], uint8_t ry[64], const uint8_t* xycs, size_t size) {
    if (!rx || !ry || (size && !xycs) || size % kEntry) return 0;
    const size_t n = size / kEntry;
    if (n >= (size_t)1 << 26) return 0;
    std::vector<uint8_t> wire(96 * n), sc(32 * n);
    for (size_t i = 0; i < n; ++i) {
        const uint8_t* e = xycs + i * kEntry;
        if (!fp_encoding_ok(e) || !fp_encoding_ok(e + 64)) return 0;
        std::memcpy(&wire[96 * i], e + 16, 48);
        std::memcpy(&wire[96 * i + 48], e + 80, 48);
        for (int k = 0; k < 32; ++k) sc[32 * i + k] = e[128 + 31 - k];  // BE -> LE
    }
    if (n == 0) {
        std::memset(rx, 0, 64);
        std::memset(ry, 0, 64);
        return 1;
    }
    try {
        Engine& e = engine();
        std::lock_guard<std::mutex> g(e.mu);
        hipStream_t s = e.stream;
        uint8_t* d_sc = static_cast<uint8_t*>(e.h2d_a.ensure(32 * n));
        uint8_t* d_wire = static_cast<uint8_t*>(e.h2d_b.ensure(96 * n));
        uint32_t* d_pts = static_cast<uint32_t*>(e.pts.ensure((size_t)affine_words(2) * 4 * n));
        uint8_t* d_inf = static_cast<uint8_t*>(e.inf.ensure(n));
        uint8_t* d_out = static_cast<uint8_t*>(e.out.ensure(512));
        int* d_bad = reinterpret_cast<int*>(d_out + 256);  // past the 144-byte MSM record
        KZ_HIP(hipMemcpyAsync(d_sc, sc.data(), 32 * n, hipMemcpyHostToDevice, s));
        KZ_HIP(hipMemcpyAsync(d_wire, wire.data(), 96 * n, hipMemcpyHostToDevice, s));
        points_from_wire(2, d_wire, d_pts, d_inf, n, s);
        device_fill_u32(reinterpret_cast<uint32_t*>(d_bad), 0, 1, s);
        check_g1(2, d_pts, d_inf, n, d_bad, s);
        int bad = 0;
        KZ_HIP(hipMemcpyAsync(&bad, d_bad, sizeof bad, hipMemcpyDeviceToHost, s));
        KZ_HIP(hipStreamSynchronize(s));
        if (bad) return 0;  // off the curve or outside G1
        MsmParams prm;
        prm.g1 = true;
        msm_device(2, d_pts, d_inf, reinterpret_cast<const uint32_t*>(d_sc), n, prm, e.ws, s, d_out);
        uint8_t res[96];
        KZ_HIP(hipMemcpyAsync(res, d_out, 96, hipMemcpyDeviceToHost, s));
        KZ_HIP(hipStreamSynchronize(s));
        std::memset(rx, 0, 16);
        std::memset(ry, 0, 16);
        std::memcpy(rx + 16, res, 48);
        std::memcpy(ry + 16, res + 48, 48);
        return 1;
    } catch (const std::exception& ex) {
        set_last_error(ex.what());
        return 0;
    }
}
