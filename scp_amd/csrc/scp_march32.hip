// scp_march32.hip — FP32C instances of the LDS-staged march (built with FMA
// contraction: the compensated sums form their products with explicit fmaf).
// Default tile: one wave per row (64 x 4 sites, 71 KB of LDS, two blocks of
// four waves per CU: while one block waits at its plane barrier the other
// computes). Two waves per row (128 x 4, one block of eight waves per CU)
// measured 20 % slower at 512^3 and is kept for the tuning harness.
#include "scp_march.h"

namespace scp {

template <int LOK, int KW>
static int launch_lo(int mode, bool emit, const KParams &kp, const Bufs32 &b, int DI, hipStream_t s) {
    return mode == 2   ? launch_march_t<S32<LOK>, KW, 2>(emit, kp, b, DI, s)
           : mode == 1 ? launch_march_t<S32<LOK>, KW, 1>(emit, kp, b, DI, s)
           : mode == 3 ? launch_march_t<S32<LOK>, KW, 3>(emit, kp, b, DI, s)
                       : launch_march_t<S32<LOK>, KW, 0>(emit, kp, b, DI, s);
}

template <int KW>
static int launch_kw(int mode, bool emit, int lo_kind, const KParams &kp, const Bufs32 &b, int DI, hipStream_t s) {
    if (lo_kind == LO_F32) return launch_lo<LO_F32, KW>(mode, emit, kp, b, DI, s);
    if (lo_kind == LO_BF16) return launch_lo<LO_BF16, KW>(mode, emit, kp, b, DI, s);
    if (lo_kind == LO_Q8) return launch_lo<LO_Q8, KW>(mode, emit, kp, b, DI, s);
    if (lo_kind == LO_Q8P) return launch_lo<LO_Q8P, KW>(mode, emit, kp, b, DI, s);
    return launch_lo<LO_NONE, KW>(mode, emit, kp, b, DI, s);
}

int launch_march_f32c_kw(int kw, int mode, bool emit, int lo_kind, const KParams &kp, const Bufs32 &b, int DI,
                         hipStream_t s) {
    if (mode < 0 || mode > 3) return (int)hipErrorInvalidValue;
    return kw == 1 ? launch_kw<1>(mode, emit, lo_kind, kp, b, DI, s) : launch_kw<2>(mode, emit, lo_kind, kp, b, DI, s);
}

int launch_march_f32c(int mode, bool emit, int lo_kind, const KParams &kp, const Bufs32 &b, int DI,
                      hipStream_t s) {
    return launch_march_f32c_kw(1, mode, emit, lo_kind, kp, b, DI, s);
}

template <int LOK>
static int diag_lo(int mode, const KParams &kp, const Bufs32 &b, int DI, const DiagArgs &da, hipStream_t s) {
    return mode == 2   ? launch_march_diag_t<S32<LOK>, 2>(kp, b, DI, da, s)
           : mode == 1 ? launch_march_diag_t<S32<LOK>, 1>(kp, b, DI, da, s)
           : mode == 3 ? launch_march_diag_t<S32<LOK>, 3>(kp, b, DI, da, s)
                       : launch_march_diag_t<S32<LOK>, 0>(kp, b, DI, da, s);
}

int launch_march_f32c_diag(int mode, int lo_kind, const KParams &kp, const Bufs32 &b, int DI, const DiagArgs &da,
                           hipStream_t s) {
    if (mode < 0 || mode > 3) return (int)hipErrorInvalidValue;
    if (lo_kind == LO_F32) return diag_lo<LO_F32>(mode, kp, b, DI, da, s);
    if (lo_kind == LO_BF16) return diag_lo<LO_BF16>(mode, kp, b, DI, da, s);
    if (lo_kind == LO_Q8) return diag_lo<LO_Q8>(mode, kp, b, DI, da, s);
    if (lo_kind == LO_Q8P) return diag_lo<LO_Q8P>(mode, kp, b, DI, da, s);
    return diag_lo<LO_NONE>(mode, kp, b, DI, da, s);
}

template <int LOK, int KW, int TJ>
static int batch_lo(int mode, const BatchGrid<Bufs32> *grids, const BatchTile *tiles, uint32_t ntiles,
                    uint32_t nblocks, uint32_t per_block, int q, hipStream_t s) {
    return mode == 2   ? launch_march_batch_t<S32<LOK>, 2, KW, TJ>(grids, tiles, ntiles, nblocks, per_block, q, s)
           : mode == 1 ? launch_march_batch_t<S32<LOK>, 1, KW, TJ>(grids, tiles, ntiles, nblocks, per_block, q, s)
                       : launch_march_batch_t<S32<LOK>, 0, KW, TJ>(grids, tiles, ntiles, nblocks, per_block, q, s);
}

template <int LOK>
static int batch_mix(int mode, const BatchGrid<Bufs32> *grids, const BatchTile *tiles, uint32_t ntiles,
                     uint32_t nblocks, uint32_t per_block, int q, hipStream_t s) {
    return mode == 2   ? launch_march_batch_mix_t<S32<LOK>, 2>(grids, tiles, ntiles, nblocks, per_block, q, s)
           : mode == 1 ? launch_march_batch_mix_t<S32<LOK>, 1>(grids, tiles, ntiles, nblocks, per_block, q, s)
                       : launch_march_batch_mix_t<S32<LOK>, 0>(grids, tiles, ntiles, nblocks, per_block, q, s);
}

// shape (TILE_*): 64 x 4 for every storage kind; the other shapes for the
// production words (int8 on the positions) only
int launch_march_batch_f32c(int mode, int lo_kind, int shape, const BatchGrid<Bufs32> *grids,
                            const BatchTile *tiles, uint32_t ntiles, uint32_t nblocks, uint32_t per_block, int q,
                            hipStream_t s) {
    if (mode < 0 || mode > 2) return (int)hipErrorInvalidValue;
    if (shape != TILE_64x4) {
        if (lo_kind != LO_Q8P) return (int)hipErrorInvalidValue;
        if (shape == TILE_128x4) return batch_lo<LO_Q8P, 2, 4>(mode, grids, tiles, ntiles, nblocks, per_block, q, s);
        if (shape == TILE_64x8) return batch_lo<LO_Q8P, 1, 8>(mode, grids, tiles, ntiles, nblocks, per_block, q, s);
        if (shape == TILE_64x2) return batch_lo<LO_Q8P, 1, 2>(mode, grids, tiles, ntiles, nblocks, per_block, q, s);
        if (shape == TILE_32x8) return batch_lo<LO_Q8P, 0, 8>(mode, grids, tiles, ntiles, nblocks, per_block, q, s);
        if (shape == TILE_MIXED) return batch_mix<LO_Q8P>(mode, grids, tiles, ntiles, nblocks, per_block, q, s);
        return (int)hipErrorInvalidValue;
    }
    if (lo_kind == LO_F32) return batch_lo<LO_F32, 1, 4>(mode, grids, tiles, ntiles, nblocks, per_block, q, s);
    if (lo_kind == LO_BF16) return batch_lo<LO_BF16, 1, 4>(mode, grids, tiles, ntiles, nblocks, per_block, q, s);
    if (lo_kind == LO_Q8) return batch_lo<LO_Q8, 1, 4>(mode, grids, tiles, ntiles, nblocks, per_block, q, s);
    if (lo_kind == LO_Q8P) return batch_lo<LO_Q8P, 1, 4>(mode, grids, tiles, ntiles, nblocks, per_block, q, s);
    return batch_lo<LO_NONE, 1, 4>(mode, grids, tiles, ntiles, nblocks, per_block, q, s);
}

}  // namespace scp
