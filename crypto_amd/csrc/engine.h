// Process-wide device state behind the C ABI: one engine per process on the
// current HIP device, holding the stream, the reusable MSM workspace and the
// staging buffers for host-pointer entry points.  Entry points that share it
// serialise on its mutex, which keeps the C ABI reentrant from many threads.
#pragma once

#include <mutex>

#include "common.h"
#include "msm.h"

namespace kz {

struct Engine {
    int device = 0;
    hipStream_t stream = nullptr;
    std::mutex mu;
    MsmWorkspace ws;
    DevBuf h2d_a, h2d_b, pts, inf, out;
    DevBuf err;  // validation flags of the Banderwagon reader
    DevBuf ntt_a, ntt_b, ntt_tw;
    DevBuf gen_table[4];  // fixed-base tables, indexed by curve id
};

// Returns the engine, initialising HIP on first use; throws HipError when no
// device is usable.
Engine& engine();

// Installs the tuning file named by KINET_HIP_TUNE_FILE, if set (throws on a
// malformed file).  Called once by engine().
void tune_load_env();

}  // namespace kz
