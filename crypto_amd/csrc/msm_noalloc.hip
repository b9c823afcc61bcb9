// kinet_hip_msm_prepared_noalloc: the MSM of kinet_hip_msm_prepared for
// callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops, harnesses that forbid hipMalloc).  Its scratch -- the
// sort histograms and entries, bucket sums, carries, reduction levels -- is
// carved out of one file-scope device array, `scratch`, so a call allocates
// nothing: no hipMalloc, hipFree, HIP event or stream creation.
//
// Set-up, once per device, on the first call with that device current:
// hipGetSymbolAddress of `scratch` (each device holds its own copy of the
// array, at its own address) and one HIP event that orders successive calls.
// Host code uses only that address, never the array's name.
//
// The array keeps the last call's contents; the pipeline clears what it needs
// cleared (bucket counters) by fill kernels in every call.  Calls on one
// device are serialised on the host, and each call's stream waits for the
// previous call's event, so calls on different streams do not overlap on the
// device.  Under graph capture that wait is not recorded: the order of graph
// replays against other calls on the same device is then the caller's, and
// overlapping ones race on the array.
#include <mutex>

#include "kinet/crypto/hip_runtime.h"
#include "kinet/gpukit/gpukit.h"
#include "msm.h"
#include "scan.h"

namespace kz {
namespace {

// Largest n: 2^16 points.  Worst case need, BLS12-381 at n = 2^16 with c = 12
// and a reduction-segment override of 1 bucket: 48.3 MiB; 64 MiB leaves room.
constexpr size_t kNoAllocMaxN = (size_t)1 << 16;
constexpr size_t kScratchBytes = (size_t)64 << 20;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) uint8_t scratch[kScratchBytes];

struct NoAllocDevice {
    bool ready = false;
    uint8_t* ptr = nullptr;  // device address of `scratch` on this device
    MsmWorkspace ws;         // pieces of scratch, carved afresh per call
};

std::mutex g_mu;
// Created at a device's set-up and kept until the process ends (no static
// destructor calls into the HIP runtime at exit).
NoAllocDevice* g_dev[kMaxDevices];

__global__ void k_fill_bytes(uint8_t* __restrict__ p, uint8_t v, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Per-device set-up; called with g_mu held.
NoAllocDevice& device_state() {
    int dev = 0;
    KZ_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("device id beyond the no-alloc device table");
    if (!g_dev[dev]) g_dev[dev] = new NoAllocDevice();
    NoAllocDevice& d = *g_dev[dev];
    if (!d.ready) {
        void* p = nullptr;
        KZ_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw std::runtime_error("hipGetSymbolAddress returned no address for the MSM scratch array");
        d.ptr = static_cast<uint8_t*>(p);
        KZ_HIP(hipEventCreateWithFlags(&d.ws.last_use, hipEventDisableTiming));
        d.ready = true;
    }
    return d;
}

}  // namespace
}  // namespace kz

using namespace kz;

extern "C" int kinet_hip_msm_prepared_noalloc(uint32_t curve, const void* d_scalars, const void* d_pts,
                                              const void* d_inf, size_t n, void* d_out, void* stream) {
    if (curve > 2) return GPUKIT_ERR_BACKEND;
    if (!d_out || (n > 0 && (!d_scalars || !d_pts))) return GPUKIT_ERR_NULL_ARG;
    if (n > kNoAllocMaxN) {
        set_last_error("kinet_hip_msm_prepared_noalloc: n = " + std::to_string(n) + " exceeds the limit of " +
                       std::to_string(kNoAllocMaxN) + " points of the static scratch array");
        return GPUKIT_ERR_BAD_SIZE;
    }
    try {
        std::lock_guard<std::mutex> g(g_mu);
        NoAllocDevice& d = device_state();
        hipStream_t s = static_cast<hipStream_t>(stream);
        uint8_t* out = static_cast<uint8_t*>(d_out);
        if (n == 0) {
            const uint32_t nb = (uint32_t)msm_out_bytes((int)curve);
            hipLaunchKernelGGL(k_fill_bytes, dim3(1), dim3(256), 0, s, out, (uint8_t)0, nb);
            KZ_HIP(hipGetLastError());
            compress_on_device((int)curve, out, out + point_wire_bytes((int)curve), s);
            return GPUKIT_OK;
        }
        DevArena arena{d.ptr, kScratchBytes, 0};
        d.ws.bind(&arena);
        NoStageTimers no_events;
        MsmParams prm;
        prm.c = msm_default_window_bits(n);  // the built-in rule: scratch sized for it
        msm_device((int)curve, static_cast<const uint32_t*>(d_pts), static_cast<const uint8_t*>(d_inf),
                   static_cast<const uint32_t*>(d_scalars), n, prm, d.ws, s, out);
        d.ws.bind(nullptr);
        return GPUKIT_OK;
    } catch (const std::length_error& ex) {
        set_last_error(ex.what());
        return GPUKIT_ERR_BAD_SIZE;
    } catch (const std::exception& ex) {
        set_last_error(ex.what());
        return GPUKIT_ERR_BACKEND;
    }
}
