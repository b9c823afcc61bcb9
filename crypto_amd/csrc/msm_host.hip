// Curve dispatch of the MSM pipeline and the window-size default.
#include "msm_internal.h"

namespace kz {

// Window bits from n: about 32 points per bucket on average (2^(c-1) ~ n/32),
// clamped to the instantiated range.
int msm_default_window_bits(size_t n) {
    int lg = 0;
    while (((size_t)1 << (lg + 1)) <= n) ++lg;
    int c = lg - 4;
    if (c < 5) c = 5;
    if (c > 16) c = 16;
    return c;
}

void msm_device(int curve, const uint32_t* d_pts, const uint8_t* d_inf, const uint32_t* d_scalars, size_t n,
                const MsmParams& prm, MsmWorkspace& ws, hipStream_t s, uint8_t* d_out) {
    if (curve == 2) msm_device_bls(d_pts, d_inf, d_scalars, n, prm, ws, s, d_out);
    else msm_device_bn(d_pts, d_inf, d_scalars, n, prm, ws, s, d_out);
}

void points_from_wire(int curve, const uint8_t* d_wire, uint32_t* d_pts, uint8_t* d_inf, size_t n,
                      hipStream_t s) {
    if (curve == 2) points_from_wire_bls(d_wire, d_pts, d_inf, n, s);
    else points_from_wire_bn(d_wire, d_pts, d_inf, n, s);
}

void compress_on_device(int curve, const uint8_t* d_wire, uint8_t* d_out, hipStream_t s) {
    if (curve == 2) compress_bls(d_wire, d_out, s);
    else compress_bn(d_wire, d_out, s);
}

}  // namespace kz
