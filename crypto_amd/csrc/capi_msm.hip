// C ABI of the MSM path.
//
//   gpukit_multi_pippenger_hip  new HIP backend next to the reference's
//                               _cpu/_cuda/_metal/_wgsl entries
//                               (gpukit/include/kinet/gpukit/multi_pippenger.h:69-92),
//                               called by the reference's own dispatcher
//                               (gpukit/c-abi/c_multi_pippenger.cpp:9-42) once
//                               it has a GPUKIT_BACKEND_HIP case
//   kinet_hip_*                 device-pointer variants for callers whose data
//                               already lives in HBM (bench, batching layers)
#include <cstring>
#include <string>

#include "engine.h"
#include "glv.h"
#include "msm_stages.h"
#include "scan.h"
#include "kinet/gpukit/multi_pippenger.h"
#include "kinet/crypto/hip_runtime.h"

namespace kz {

static std::mutex g_err_mu;
static std::string g_last_error;

void set_last_error(const std::string& msg) {
    std::lock_guard<std::mutex> g(g_err_mu);
    g_last_error = msg;
}

Engine& engine() {
    static Engine* e = [] {
        int count = 0;
        hipError_t rc = hipGetDeviceCount(&count);
        if (rc != hipSuccess || count == 0) throw HipError(hipErrorNoDevice, "no HIP device available");
        Engine* eng = new Engine();
        KZ_HIP(hipGetDevice(&eng->device));
        KZ_HIP(hipStreamCreateWithFlags(&eng->stream, hipStreamNonBlocking));
        tune_load_env();  // tuning tables named by KINET_HIP_TUNE_FILE
        return eng;
    }();
    return *e;
}

// Curves of the device entries: BN254 and BLS12-381 G1 everywhere; secp256k1
// through the plain MSM and the GLV form, not through window tables;
// Banderwagon through the plain MSM only (gpukit_multi_pippenger_hip,
// points_prepare, msm_prepared).
static bool valid_curve(uint32_t curve) { return curve == 1 || curve == 2; }
static bool valid_glv_curve(uint32_t curve) { return curve <= 2; }
static bool valid_plain_curve(uint32_t curve) { return curve <= 3; }

// Banderwagon's validating reader on stream s; returns the error flags
// (bit 0: a point the reference's deserialize_uncompressed rejects, bit 1: a
// scalar not below the group order), synchronising s.  d_sc may be null.
static uint32_t bw_read(Engine& e, const uint8_t* d_wire, const uint32_t* d_sc, size_t n, uint32_t* d_pts,
                        uint8_t* d_inf, hipStream_t s) {
    uint32_t* d_err = static_cast<uint32_t*>(e.err.ensure(sizeof(uint32_t)));
    device_fill_u32(d_err, 0, 1, s);
    bw_points_from_wire(d_wire, d_pts, d_inf, (uint32_t)n, d_err, s);
    if (d_sc) bw_check_scalars(d_sc, (uint32_t)n, d_err, s);
    uint32_t err = 0;
    KZ_HIP(hipMemcpyAsync(&err, d_err, sizeof err, hipMemcpyDeviceToHost, s));
    KZ_HIP(hipStreamSynchronize(s));
    return err;
}

}  // namespace kz

using namespace kz;

extern "C" {

const char* kinet_hip_last_error(void) {
    std::lock_guard<std::mutex> g(g_err_mu);
    return g_last_error.c_str();
}

int kinet_hip_available(void) {
    try {
        (void)engine();
        return 1;
    } catch (const std::exception& ex) {
        set_last_error(ex.what());
        return 0;
    }
}

int gpukit_multi_pippenger_hip(uint32_t curve, const uint8_t* scalars, const uint8_t* points, size_t n,
                               uint8_t* result) {
    if (!result) return GPUKIT_ERR_NULL_ARG;
    if (n > 0 && (!scalars || !points)) return GPUKIT_ERR_NULL_ARG;
    if (!valid_plain_curve(curve)) return GPUKIT_ERR_BACKEND;
    const size_t pb = point_wire_bytes((int)curve);
    if (n == 0) {
        std::memset(result, 0, pb);
        if (curve == GPUKIT_CURVE_BANDERWAGON) result[63] = 1;  // the identity (0, 1)
        return GPUKIT_OK;
    }
    if (n >= (size_t)1 << 32) return GPUKIT_ERR_BAD_SIZE;  // 32-bit point indices
    try {
        Engine& e = engine();
        std::lock_guard<std::mutex> g(e.mu);
        hipStream_t s = e.stream;
        uint8_t* d_sc = static_cast<uint8_t*>(e.h2d_a.ensure(32 * n));
        uint8_t* d_wire = static_cast<uint8_t*>(e.h2d_b.ensure(pb * n));
        uint32_t* d_pts = static_cast<uint32_t*>(e.pts.ensure((size_t)affine_words((int)curve) * 4 * n));
        uint8_t* d_inf = static_cast<uint8_t*>(e.inf.ensure(n));
        uint8_t* d_out = static_cast<uint8_t*>(e.out.ensure(msm_out_bytes((int)curve)));
        KZ_HIP(hipMemcpyAsync(d_sc, scalars, 32 * n, hipMemcpyHostToDevice, s));
        KZ_HIP(hipMemcpyAsync(d_wire, points, pb * n, hipMemcpyHostToDevice, s));
        if (curve == GPUKIT_CURVE_BANDERWAGON) {
            // the reference rejects the whole call for any bad point or scalar
            // (multi_pippenger.cpp:324-339)
            if (bw_read(e, d_wire, reinterpret_cast<const uint32_t*>(d_sc), n, d_pts, d_inf, s))
                return GPUKIT_ERR_BAD_SIZE;
        } else {
            points_from_wire((int)curve, d_wire, d_pts, d_inf, n, s);
        }
        MsmParams prm;
        // BN254 G1 and secp256k1 have cofactor 1: once every point is on the
        // curve, the GLV split applies (the same bytes, fewer windows).
        // Off-curve input keeps the plain path.  BLS12-381 would need the
        // subgroup check, which costs more than the split saves in one call.
        if ((curve == GPUKIT_CURVE_BN254_G1 || curve == GPUKIT_CURVE_SECP256K1) && n >= 4096 &&
            msm_glv_enabled((int)curve, n)) {
            int* d_bad = static_cast<int*>(e.err.ensure(sizeof(int)));
            device_fill_u32(reinterpret_cast<uint32_t*>(d_bad), 0, 1, s);
            check_g1((int)curve, d_pts, d_inf, n, d_bad, s);
            int bad = 1;
            KZ_HIP(hipMemcpyAsync(&bad, d_bad, sizeof bad, hipMemcpyDeviceToHost, s));
            KZ_HIP(hipStreamSynchronize(s));
            prm.g1 = bad == 0;
        }
        msm_device((int)curve, d_pts, d_inf, reinterpret_cast<const uint32_t*>(d_sc), n, prm, e.ws, s, d_out);
        KZ_HIP(hipMemcpyAsync(result, d_out, pb, hipMemcpyDeviceToHost, s));
        KZ_HIP(hipStreamSynchronize(s));
        return GPUKIT_OK;
    } catch (const std::exception& ex) {
        set_last_error(ex.what());
        return GPUKIT_ERR_BACKEND;
    }
}

int kinet_hip_points_prepare(uint32_t curve, const void* d_wire, size_t n, void* d_pts, void* d_inf,
                             void* stream) {
    if (!valid_plain_curve(curve)) return GPUKIT_ERR_BACKEND;
    if (n > 0 && (!d_wire || !d_pts || !d_inf)) return GPUKIT_ERR_NULL_ARG;
    if (n >= (size_t)1 << 32) return GPUKIT_ERR_BAD_SIZE;
    try {
        Engine& e = engine();
        hipStream_t s = static_cast<hipStream_t>(stream);
        if (curve == GPUKIT_CURVE_BANDERWAGON) {
            std::lock_guard<std::mutex> g(e.mu);
            return bw_read(e, static_cast<const uint8_t*>(d_wire), nullptr, n, static_cast<uint32_t*>(d_pts),
                           static_cast<uint8_t*>(d_inf), s)
                       ? GPUKIT_ERR_BAD_SIZE
                       : GPUKIT_OK;
        }
        points_from_wire((int)curve, static_cast<const uint8_t*>(d_wire), static_cast<uint32_t*>(d_pts),
                         static_cast<uint8_t*>(d_inf), n, s);
        return GPUKIT_OK;
    } catch (const std::exception& ex) {
        set_last_error(ex.what());
        return GPUKIT_ERR_BACKEND;
    }
}

static int msm_prepared_impl(uint32_t curve, const void* d_scalars, const void* d_pts, const void* d_inf,
                             size_t n, int window_bits, int chunk, void* d_out, void* stream, bool g1) {
    if (g1 ? !valid_glv_curve(curve) : !valid_plain_curve(curve)) return GPUKIT_ERR_BACKEND;
    if (!d_out || (n > 0 && (!d_scalars || !d_pts))) return GPUKIT_ERR_NULL_ARG;
    if (window_bits != 0 && (window_bits < 5 || window_bits > 18)) return GPUKIT_ERR_BAD_SIZE;
    if (chunk < 0 || chunk > 4096) return GPUKIT_ERR_BAD_SIZE;
    if (n >= (size_t)1 << 32) return GPUKIT_ERR_BAD_SIZE;
    try {
        Engine& e = engine();
        std::lock_guard<std::mutex> g(e.mu);
        hipStream_t s = static_cast<hipStream_t>(stream);
        if (n == 0) {
            if (curve == GPUKIT_CURVE_BANDERWAGON) {
                bw_identity_record(static_cast<uint8_t*>(d_out), s);
                return GPUKIT_OK;
            }
            KZ_HIP(hipMemsetAsync(d_out, 0, msm_out_bytes((int)curve), s));
            uint8_t* d_zero = static_cast<uint8_t*>(d_out);
            compress_on_device((int)curve, d_zero, d_zero + point_wire_bytes((int)curve), s);
            return GPUKIT_OK;
        }
        MsmParams prm;
        prm.c = window_bits;
        prm.g1 = g1;
        if (chunk > 0) prm.chunk = chunk;
        msm_device((int)curve, static_cast<const uint32_t*>(d_pts), static_cast<const uint8_t*>(d_inf),
                   static_cast<const uint32_t*>(d_scalars), n, prm, e.ws, s, static_cast<uint8_t*>(d_out));
        return GPUKIT_OK;
    } catch (const std::exception& ex) {
        set_last_error(ex.what());
        return GPUKIT_ERR_BACKEND;
    }
}

int kinet_hip_msm_prepared(uint32_t curve, const void* d_scalars, const void* d_pts, const void* d_inf,
                           size_t n, int window_bits, int chunk, void* d_out, void* stream) {
    return msm_prepared_impl(curve, d_scalars, d_pts, d_inf, n, window_bits, chunk, d_out, stream, false);
}

int kinet_hip_msm_prepared_g1(uint32_t curve, const void* d_scalars, const void* d_pts, const void* d_inf,
                              size_t n, int window_bits, int chunk, void* d_out, void* stream) {
    return msm_prepared_impl(curve, d_scalars, d_pts, d_inf, n, window_bits, chunk, d_out, stream, true);
}

size_t kinet_hip_points_bytes(uint32_t curve, size_t n) {
    if (!valid_plain_curve(curve)) return 0;
    return (size_t)affine_words((int)curve) * sizeof(uint32_t) * n;
}

int kinet_hip_table_window_bits(size_t n) { return msm_table_window_bits(n); }

size_t kinet_hip_table_bytes(uint32_t curve, size_t n, int window_bits) {
    if (!valid_curve(curve) || window_bits < 5 || window_bits > 18) return 0;
    return (size_t)msm_table_rows(window_bits) * n * affine_words((int)curve) * sizeof(uint32_t);
}

int kinet_hip_points_table_build(uint32_t curve, const void* d_pts, const void* d_inf, size_t n, int window_bits,
                                 void* d_table, void* stream) {
    if (!valid_curve(curve)) return GPUKIT_ERR_BACKEND;
    if (n > 0 && (!d_pts || !d_table)) return GPUKIT_ERR_NULL_ARG;
    if (window_bits < 5 || window_bits > 18) return GPUKIT_ERR_BAD_SIZE;
    try {
        (void)engine();
        points_table_build((int)curve, static_cast<const uint32_t*>(d_pts), static_cast<const uint8_t*>(d_inf), n,
                           window_bits, static_cast<uint32_t*>(d_table), static_cast<hipStream_t>(stream));
        return GPUKIT_OK;
    } catch (const std::invalid_argument& ex) {
        set_last_error(ex.what());
        return GPUKIT_ERR_BAD_SIZE;
    } catch (const std::exception& ex) {
        set_last_error(ex.what());
        return GPUKIT_ERR_BACKEND;
    }
}

int kinet_hip_msm_table(uint32_t curve, const void* d_scalars, const void* d_table, const void* d_inf,
                        size_t table_n, size_t n, int window_bits, void* d_out, void* stream) {
    if (!valid_curve(curve)) return GPUKIT_ERR_BACKEND;
    if (!d_out || (n > 0 && (!d_scalars || !d_table))) return GPUKIT_ERR_NULL_ARG;
    if (window_bits < 5 || window_bits > 18 || n > table_n || table_n >= ((size_t)1 << 31))
        return GPUKIT_ERR_BAD_SIZE;
    if (n >= (size_t)1 << 26) return GPUKIT_ERR_BAD_SIZE;
    try {
        Engine& e = engine();
        std::lock_guard<std::mutex> g(e.mu);
        hipStream_t s = static_cast<hipStream_t>(stream);
        if (n == 0) {
            KZ_HIP(hipMemsetAsync(d_out, 0, msm_out_bytes((int)curve), s));
            uint8_t* d_zero = static_cast<uint8_t*>(d_out);
            compress_on_device((int)curve, d_zero, d_zero + point_wire_bytes((int)curve), s);
            return GPUKIT_OK;
        }
        MsmParams prm;
        prm.c = window_bits;
        prm.table_stride = (uint32_t)table_n;
        msm_device((int)curve, static_cast<const uint32_t*>(d_table), static_cast<const uint8_t*>(d_inf),
                   static_cast<const uint32_t*>(d_scalars), n, prm, e.ws, s, static_cast<uint8_t*>(d_out));
        return GPUKIT_OK;
    } catch (const std::invalid_argument& ex) {
        set_last_error(ex.what());
        return GPUKIT_ERR_BAD_SIZE;
    } catch (const std::exception& ex) {
        set_last_error(ex.what());
        return GPUKIT_ERR_BACKEND;
    }
}

void kinet_hip_profile_enable(int on) { msm_set_profiling(on); }

int kinet_hip_msm_set_reduce_segment(int seg) {
    try {
        msm_set_reduce_segment(seg);
        return GPUKIT_OK;
    } catch (const std::exception& ex) {
        set_last_error(ex.what());
        return GPUKIT_ERR_BAD_SIZE;
    }
}

int kinet_hip_profile_collect(double* stage_ms, int* stage_calls, int nstages) {
    if (!stage_ms || !stage_calls) return GPUKIT_ERR_NULL_ARG;
    try {
        return msm_collect_profile(stage_ms, stage_calls, nstages);
    } catch (const std::exception& ex) {
        set_last_error(ex.what());
        return GPUKIT_ERR_BACKEND;
    }
}

}  // extern "C"
