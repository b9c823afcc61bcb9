// Host-side helpers shared by the HIP library: status codes, error checks and
// a grow-only device buffer.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <cstdio>
#include <mutex>
#include <stdexcept>
#include <string>

// gpukit status codes (gpukit/include/kinet/gpukit/gpukit.h:38-46).
#define KZ_GPUKIT_OK 0
#define KZ_GPUKIT_ERR_NULL_ARG (-1)
#define KZ_GPUKIT_ERR_BAD_SIZE (-2)
#define KZ_GPUKIT_ERR_BACKEND (-3)
#define KZ_GPUKIT_ERR_DISPATCH (-4)
#define KZ_GPUKIT_ERR_NOTIMPL (-5)

// crypto status codes (c-abi/crypto.h:60-66).
#define KZ_CRYPTO_OK 0
#define KZ_CRYPTO_ERR_INPUT (-1)
#define KZ_CRYPTO_ERR_LENGTH (-2)
#define KZ_CRYPTO_ERR_VERIFY (-3)
#define KZ_CRYPTO_ERR_BACKEND (-4)
#define KZ_CRYPTO_ERR_NOTIMPL (-5)
#define KZ_CRYPTO_ERR_INTERNAL (-6)

namespace kz {

struct HipError : std::runtime_error {
    hipError_t code;
    HipError(hipError_t c, const char* what) : std::runtime_error(what), code(c) {}
};

inline void hip_check(hipError_t e, const char* expr, const char* file, int line) {
    if (e != hipSuccess) {
        char buf[512];
        std::snprintf(buf, sizeof buf, "HIP error %d (%s) at %s:%d: %s", (int)e, hipGetErrorString(e), file,
                      line, expr);
        throw HipError(e, buf);
    }
}

#define KZ_HIP(expr) ::kz::hip_check((expr), #expr, __FILE__, __LINE__)

// Records the last error message for kinet_hip_last_error().
void set_last_error(const std::string& msg);

// Bump allocator over device memory the library already holds (a static
// device array): hands out 256-byte aligned pieces and never calls hipMalloc.
// Running past the end throws std::length_error before anything is launched.
struct DevArena {
    uint8_t* base = nullptr;
    size_t cap = 0, used = 0;
    void* take(size_t bytes) {
        const size_t off = (used + 255) & ~(size_t)255;
        if (off > cap || bytes > cap - off)
            throw std::length_error("MSM scratch exhausted: the call needs more than the static scratch array holds");
        used = off + bytes;
        return base + off;
    }
};

// Grow-only device allocation; with an arena bound, pieces of the arena
// instead (no hipMalloc / hipFree).
struct DevBuf {
    void* ptr = nullptr;
    size_t cap = 0;
    DevArena* arena = nullptr;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() {
        if (ptr && !arena) (void)hipFree(ptr);
    }
    // Points this buffer at an arena, forgetting any earlier piece.
    void bind(DevArena* a) {
        if (ptr && !arena) (void)hipFree(ptr);
        arena = a;
        ptr = nullptr;
        cap = 0;
    }
    void* ensure(size_t bytes) {
        if (bytes > cap && arena) {
            ptr = arena->take(bytes < 256 ? 256 : bytes);
            cap = bytes < 256 ? 256 : bytes;
            return ptr;
        }
        if (bytes > cap) {
            if (ptr) KZ_HIP(hipFree(ptr));
            ptr = nullptr;
            cap = 0;
            size_t want = bytes < 256 ? 256 : bytes;
            KZ_HIP(hipMalloc(&ptr, want));
            cap = want;
        }
        return ptr;
    }
    template <class T>
    T* as() const {
        return static_cast<T*>(ptr);
    }
};

inline unsigned ceil_div(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

}  // namespace kz
