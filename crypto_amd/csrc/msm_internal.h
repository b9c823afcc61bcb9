// Internal host-side declarations shared by the MSM translation units.
#pragma once

#include <stdexcept>

#include "msm.h"

namespace kz {

// Signed-digit count (scatter = false) or scatter (scatter = true) pass.
void digits_dispatch(int c, bool scatter, const uint32_t* d_scalars, const uint8_t* d_inf, uint32_t n,
                     uint32_t* counts_or_cursor, uint32_t* entries, hipStream_t s);

// The same passes over every job of a batch (d_jobs: njobs descriptors on the
// device, nblocks = total 256-thread blocks).
void digits_multi_dispatch(bool scatter, const uint32_t* d_scalars, const uint8_t* d_inf, const MsmJob* d_jobs,
                           uint32_t njobs, uint32_t nblocks, uint32_t* counts_or_cursor, uint32_t* entries,
                           hipStream_t s);

}  // namespace kz
