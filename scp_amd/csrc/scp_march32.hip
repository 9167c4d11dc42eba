// scp_march32.hip — FP32C instances of the LDS-staged march (built with FMA
// contraction: the compensated sums form their products with explicit fmaf).
#include "scp_march.h"

namespace scp {

int launch_march_f32c(int mode, bool emit, int lo_kind, const KParams &kp, const Bufs32 &b, int DI,
                      hipStream_t s) {
    if (mode != 1 && mode != 2) return (int)hipErrorInvalidValue;
    if (lo_kind == LO_F32)
        return mode == 2 ? launch_march_t<S32<LO_F32>, 2>(emit, kp, b, DI, s)
                         : launch_march_t<S32<LO_F32>, 1>(emit, kp, b, DI, s);
    if (lo_kind == LO_BF16)
        return mode == 2 ? launch_march_t<S32<LO_BF16>, 2>(emit, kp, b, DI, s)
                         : launch_march_t<S32<LO_BF16>, 1>(emit, kp, b, DI, s);
    return mode == 2 ? launch_march_t<S32<LO_NONE>, 2>(emit, kp, b, DI, s)
                     : launch_march_t<S32<LO_NONE>, 1>(emit, kp, b, DI, s);
}

}  // namespace scp
