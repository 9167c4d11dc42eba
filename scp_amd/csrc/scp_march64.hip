// scp_march64.hip — FP64 instances of the LDS-staged march (no FMA contraction:
// the FP64 step rounds exactly like the reference CPU path). One wave per row
// (64 x 4 tiles): three staged FP64 planes of 15 fields use 142 KB of LDS
// (gauged; 6 fields, 57 KB, in real mode).
#include "scp_march.h"

namespace scp {

int launch_march_f64(int mode, bool emit, const KParams &kp, const Bufs<double> &b, int DI, hipStream_t s) {
    if (mode == 2) return launch_march_t<S64, 1, 2>(emit, kp, b, DI, s);
    if (mode == 1) return launch_march_t<S64, 1, 1>(emit, kp, b, DI, s);
    if (mode == 0) return launch_march_t<S64, 1, 0>(emit, kp, b, DI, s);
    if (mode == 3) return launch_march_t<S64, 1, 3>(emit, kp, b, DI, s);
    return (int)hipErrorInvalidValue;
}

int launch_march_f64_diag(int mode, const KParams &kp, const Bufs<double> &b, int DI, const DiagArgs &da,
                          hipStream_t s) {
    if (mode == 2) return launch_march_diag_t<S64D, 2>(kp, b, DI, da, s);
    if (mode == 1) return launch_march_diag_t<S64D, 1>(kp, b, DI, da, s);
    if (mode == 0) return launch_march_diag_t<S64D, 0>(kp, b, DI, da, s);
    if (mode == 3) return launch_march_diag_t<S64D, 3>(kp, b, DI, da, s);
    return (int)hipErrorInvalidValue;
}

int launch_march_batch_f64(int mode, const BatchGrid<Bufs<double>> *grids, const BatchTile *tiles,
                           uint32_t ntiles, uint32_t nblocks, uint32_t per_block, int q, hipStream_t s) {
    if (mode == 2) return launch_march_batch_t<S64, 2>(grids, tiles, ntiles, nblocks, per_block, q, s);
    if (mode == 1) return launch_march_batch_t<S64, 1>(grids, tiles, ntiles, nblocks, per_block, q, s);
    if (mode == 0) return launch_march_batch_t<S64, 0>(grids, tiles, ntiles, nblocks, per_block, q, s);
    return (int)hipErrorInvalidValue;
}

}  // namespace scp
