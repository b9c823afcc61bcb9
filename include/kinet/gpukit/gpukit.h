/*
 * gpukit status codes and backend ids, as in the reference
 * (gpukit/include/kinet/gpukit/gpukit.h:29-46), with the HIP backend added.
 */
#pragma once
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum gpukit_backend {
    GPUKIT_BACKEND_CPU = 0,
    GPUKIT_BACKEND_METAL = 1,
    GPUKIT_BACKEND_CUDA = 2,
    GPUKIT_BACKEND_WGSL = 3,
    GPUKIT_BACKEND_HIP = 4, /* new: AMD MI355X (gfx950), selected by GPUKIT_BACKEND=hip */
} gpukit_backend;

typedef enum gpukit_status {
    GPUKIT_OK = 0,
    GPUKIT_ERR_NULL_ARG = -1,
    GPUKIT_ERR_BAD_SIZE = -2,
    GPUKIT_ERR_BACKEND = -3,
    GPUKIT_ERR_DISPATCH = -4,
    GPUKIT_ERR_NOTIMPL = -5,
    GPUKIT_ERR_DETERMINISM = -6,
} gpukit_status;

#ifdef __cplusplus
} /* extern "C" */
#endif
