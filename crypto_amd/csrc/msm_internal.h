// Internal host-side declarations shared by the MSM translation units.
#pragma once

#include <stdexcept>

#include "msm.h"

namespace kz {

// Signed-digit count (scatter = false) or scatter (scatter = true) pass.
void digits_dispatch(int c, bool scatter, const uint32_t* d_scalars, const uint8_t* d_inf, uint32_t n,
                     uint32_t* counts_or_cursor, uint32_t* entries, hipStream_t s);

}  // namespace kz
