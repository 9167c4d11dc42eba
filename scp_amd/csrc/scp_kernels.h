// scp_kernels.h — internal interface between the HIP kernels (scp_step.hip)
// and the C-ABI layer (scp_abi.hip). Not installed; include/scp_amd.h is the
// public boundary.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace scp {

// Stencil ("position") fields on the device, per mode:
//   real    NP=6 : u0..2, tu0..2
//   complex NP=12: u0..2, v0..2, tu0..2, tv0..2
//   gauged  NP=15: u0..2, v0..2, tu0..2, tv0..2, th0..2
// Velocity-like fields use the same numbering (ud, vd, tud, tvd, E).
enum { F_U = 0, F_V = 3, F_TU = 6, F_TV = 9, F_TH = 12 };

inline int np_of_mode(int mode) { return mode == 2 ? 15 : (mode == 1 ? 12 : 6); }

// Physics constants, derived on the host with the reference's own expressions
// (scp_sim.c:204-205, 746-747, 1057-1063, 1642) so they round identically.
struct KParams {
    int32_t N, rmode;
    uint32_t n3, fs;           // sites, field stride (elements, 64-aligned)
    double hdt, dt, gad;       // 0.5*dt, dt, -g*dx*dt
    double h1, h2, inva;       // 1/(2dx), 1/dx^2, 1/dx
    double STP, GG;            // 1/(g dx^3), g
    double MU, KAPPA, MASS2, MT2, ETA;
    double KH, ACS, BH, IA, IB, KG;   // real-mode optional terms
    // real-mode terms of the reference's GPU step (scp_sim.cu:1144-1248); XT: any of them on
    double TSAT, GCONV, SGRAD, CHI, SCUB, SFREQ, SCROSS, LSELF;
    int32_t XT, pad0;
    double L, dx, Rd, DW, DR;  // sponge: R_damp = L - damp_width
    int32_t damp;              // 1 when the sponge is active (bc 0 with width, rate > 0)
    int32_t bc1;               // 1: gradient-pinned boundary (bc 1), x-slabs of gm planes per face
    // Layout of the step state P, W and their compensation words (the arrays
    // the step streams): rows padded to 32 elements (128 B in FP32), so every
    // 64-site row segment of a field starts on a cache-line boundary whatever
    // N is. The emitted state xo/vo and the diagnostics arrays keep the
    // unpadded layout i*N*N + j*N + k with field stride fs.
    uint32_t pitch;            // row pitch (elements) of P/W: N rounded up to 32 (N in the real mode)
    uint32_t plane;            // N * pitch
    uint32_t fsh;              // field stride of P/W: N * plane rounded up to 64
    int32_t gm;                // bc 1: gradient_margin
    int32_t init_acc;          // K_INIT: the first half kick takes acc (a given a_n), not F(x)
};

// Device buffers. Each group is one allocation of nfields * fs elements.
template <typename T>
struct Bufs {
    const T *pin;   // positions at the stencil time level (read with neighbours)
    T *pout;        // positions one drift later
    T *w;           // half-kicked velocities (read and written at the own site)
    T *wlo;         // compensation words of w (FP32C only)
    const T *pinlo; // compensation words of pin (FP32C only)
    T *poutlo;
    double *xo;     // emitted positions (FP64), also the INIT source
    double *vo;     // emitted velocities (FP64), also the INIT source
    double *acc;    // forces-only output (FP64)
    const double *aux;  // real mode: Cosserat mismatch M[3], hardening Q[3], theta frequency (7 fields)
    // gradient-pinned boundary (bc 1, FP64): pinned positions / velocities
    // (emitted layout) and the side rows the y-edge pass reads, per field f:
    // side[(f*6 + r)*N*N + i*N + k], r = 0/1 acceleration at j = 0 / N-1,
    // r = 2..5 velocity at j = 1, 2, N-2, N-3
    const double *pinx, *pinv;
    double *side;
};

enum StepKind { K_INIT = 0, K_STEP = 1, K_FORCES = 2 };

// FP32 compensated state: hi words (FP32) plus a compensation ("lo") word per
// scalar whose storage is selected by LoKind.
enum LoKind { LO_F32 = 0, LO_BF16 = 1, LO_NONE = 2, LO_Q8 = 3, LO_Q8P = 4 };
// LO_Q8P: LO_Q8 words on the positions P only, velocities W without
constexpr int lo_kind_w(int k) { return k == LO_Q8P ? LO_NONE : k; }   // storage of W's words
constexpr int lo_kind_p(int k) { return k == LO_Q8P ? LO_Q8 : k; }     // storage of P's words

// Storage type and codec of a compensation word. The lo word of a pair holds
// at most ulp(hi)/2 in magnitude (the exact error of the TwoSum that formed hi).
//   LO_F32  the FP32 value;
//   LO_BF16 the FP32 value rounded to bfloat16 (nearest even);
//   LO_Q8   a signed byte q in units of ulp(hi)/256: the byte keeps what a
//           bfloat16 keeps of the largest lo words in half the storage;
//   LO_NONE nothing (plain FP32).
template <int LOK> struct LoWord;
template <> struct LoWord<LO_F32> {
    using W = float;
    static constexpr bool STORED = true;
    __device__ static float dec(W w, float) { return w; }
    __device__ static W enc(float lo, float) { return lo; }
};
template <> struct LoWord<LO_BF16> {
    using W = uint16_t;
    static constexpr bool STORED = true;
    __device__ static float dec(W w, float) { return __uint_as_float((uint32_t)w << 16); }
    __device__ static W enc(float lo, float) {     // round to nearest even
        uint32_t u = __float_as_uint(lo);
        u += 0x7FFFu + ((u >> 16) & 1u);
        return (W)(u >> 16);
    }
};
template <> struct LoWord<LO_Q8> {
    using W = int8_t;
    static constexpr bool STORED = true;
    // lo = q * 2^(E - 32) with hi = m * 2^E, 1/2 <= |m| < 1 (the exponent frexp
    // gives): 256 steps per ulp(hi) = 2^(E - 24), exact scaling by v_ldexp
    __device__ static float dec(W w, float hi) {
        return __builtin_amdgcn_ldexpf((float)w, __builtin_amdgcn_frexp_expf(hi) - 32);
    }
    __device__ static W enc(float lo, float hi) {   // round to nearest, |q| <= 127
        const int q = (int)rintf(__builtin_amdgcn_ldexpf(lo, 32 - __builtin_amdgcn_frexp_expf(hi)));
        return (W)min(max(q, -127), 127);
    }
};
template <> struct LoWord<LO_NONE> {
    using W = uint8_t;
    static constexpr bool STORED = false;
    __device__ static float dec(W, float) { return 0.0f; }
    __device__ static W enc(float, float) { return 0; }
};
// LO_Q8 words of one site are stored four to a 32-bit word, field f in byte
// f % 4 of word f / 4 (a word array per four fields, laid out like a field):
// a wave then stores whole 256-byte runs, never a part of a cache line that a
// block on another XCD completes.
struct LoQuads {
    uint32_t q[4];
    __device__ int8_t get(int f) const { return (int8_t)(q[f >> 2] >> (8 * (f & 3))); }
    __device__ void set(int f, int8_t v) { q[f >> 2] |= (uint32_t)(uint8_t)v << (8 * (f & 3)); }
};
inline int lo_quads(int nf) { return (nf + 3) / 4; }
// bytes of compensation words per site of one NP-field buffer
inline int lo_site_bytes(int lo_kind, int np) {
    return lo_kind == LO_F32 ? 4 * np : lo_kind == LO_BF16 ? 2 * np : lo_kind == LO_Q8 ? 4 * lo_quads(np) : 0;
}
struct Bufs32 {
    const float *pin;    // hi words of P at the stencil level
    float *pout;
    float *w;
    const void *pinlo;   // lo words (LoKind storage) of pin / pout / w
    void *poutlo;
    void *wlo;
    double *xo, *vo;     // emitted FP64 state (hi + lo), INIT source
    double *acc;
    const double *aux;
};

// Batched march (many independent grids of one mode and storage, one launch
// per step): per-grid descriptor in device memory and a tile table of
// (grid, local tile) pairs, local tile = tk + ntk * (tj + ntj * seg).
template <class B>
struct BatchGrid {
    KParams kp;
    B buf[2];          // the grid's buffers when its P parity is 0 / 1
    int32_t DI;        // planes per segment
    int32_t cur0;      // P parity at chain launch 0
    int32_t nq;        // chain launches this grid takes part in; it emits on the last ...
    int32_t emit;      // ... when this is 1 (the grid holds its FP64 emit buffers)
    uint32_t ntk, ntj; // tiles per plane along k and j
    int32_t tk;        // tile width of this grid's tiles (sites along k; the mixed-shape launch reads it)
    int32_t pad_;
};
struct BatchTile {
    uint32_t grid, tile;
};

// Tricubic patch fits of the real-mode vector frames (scp_vec.hip): the six
// fields of the emitted state xo, patches of bs^3 (4 <= bs <= 16), 384 floats
// per patch in the reference's order (field, then a*16 + b*4 + c).
int launch_vec_fit(const double *xo, uint32_t fs, int N, int bs, float *out, hipStream_t s);

// Launchers (scp_step.hip, scp_step32.hip). Return hipError_t as int.
int launch_step_f64(int mode, int kind, bool emit, const KParams &kp, const Bufs<double> &b,
                    hipStream_t s);
int launch_step_f32c(int mode, int kind, bool emit, int lo_kind, const KParams &kp, const Bufs32 &b,
                     hipStream_t s);
// LDS-staged march (scp_march.hip): K_STEP of the complex (1) and gauged (2)
// modes; DI <= 0 picks the segment depth.
int launch_march_f64(int mode, bool emit, const KParams &kp, const Bufs<double> &b, int DI, hipStream_t s);
int launch_march_f32c(int mode, bool emit, int lo_kind, const KParams &kp, const Bufs32 &b, int DI,
                      hipStream_t s);
// the same with the tile width chosen: kw waves (64 sites each) per tile row
int launch_march_f32c_kw(int kw, int mode, bool emit, int lo_kind, const KParams &kp, const Bufs32 &b, int DI,
                         hipStream_t s);
// One batched launch, chain step q: nblocks blocks, block b runs tiles
// [b*per_block, (b+1)*per_block) of the table (XCD-aware block order).
int launch_march_batch_f64(int mode, const BatchGrid<Bufs<double>> *grids, const BatchTile *tiles,
                           uint32_t ntiles, uint32_t nblocks, uint32_t per_block, int q, hipStream_t s);
int launch_march_batch_f32c(int mode, int lo_kind, int shape, const BatchGrid<Bufs32> *grids,
                            const BatchTile *tiles, uint32_t ntiles, uint32_t nblocks, uint32_t per_block, int q,
                            hipStream_t s);
// sites along k of a march tile with kw waves per row (kw = 0: 32-site rows)
inline int march_tkm(int kw) { return kw == 0 ? 32 : 64 * kw; }
// tiles per plane of the march at N (k, j), for building tile tables
inline uint32_t march_tiles_k(int N, int kw = 1) {
    return (uint32_t)(N + march_tkm(kw) - 1) / (uint32_t)march_tkm(kw);
}
inline uint32_t march_tiles_j(int N, int tj = 4) { return (uint32_t)(N + tj - 1) / (uint32_t)tj; }
// tile shapes of the batched march: 64 x 4 (every class), and for FP32C grids
// with int8 position words 128 x 4, 64 x 8, 64 x 2 and 32 x 8 (rows of 32
// sites, two per wave: edges that are not multiples of 64 waste fewer lanes)
// TILE_MIXED: each grid its own of 64 x 4 and 32 x 8, in one launch (the grid
// shape is then given by grid_tile_shape)
enum { TILE_64x4 = 0, TILE_128x4 = 1, TILE_64x8 = 2, TILE_64x2 = 3, TILE_32x8 = 4, TILE_MIXED = 5, TILE_NSHAPES = 6 };
inline int tile_kw(int shape) { return shape == TILE_128x4 ? 2 : (shape == TILE_32x8 ? 0 : 1); }
inline int tile_tk(int shape) { return march_tkm(tile_kw(shape)); }
inline int tile_tj(int shape) { return (shape == TILE_64x8 || shape == TILE_32x8) ? 8 : (shape == TILE_64x2 ? 2 : 4); }
// resident blocks per CU of a shape (its staged planes in the 160 KB of LDS)
inline int tile_blocks_per_cu(int shape, bool f32) {
    if (!f32) return 1;
    return (shape == TILE_64x4 || shape == TILE_32x8 || shape == TILE_MIXED) ? 2 : (shape == TILE_64x2 ? 3 : 1);
}
// lanes a tile shape issues per step on an N^3 grid (tiles padded to whole tiles)
inline uint64_t tile_lanes(int shape, int N) {
    return (uint64_t)march_tiles_k(N, tile_kw(shape)) * (uint64_t)tile_tk(shape) * march_tiles_j(N, tile_tj(shape)) *
           (uint64_t)tile_tj(shape) * (uint64_t)N;
}
// the 256-thread shape an N^3 grid runs fastest, from batches of equal grids:
// 32 x 8 where it issues at most 90 % of the lanes of 64 x 4 (+5 to +62 % per
// site at edges 32, 80, 96, 144, 160, 200), or no more lanes at edges up to 128
// (+1 to +2 % at 48, 64, 128); 64 x 4 otherwise (32 x 8 0-2 % slower at 176 to
// 256 and at 512)
inline int grid_tile_shape(int N) {
    const uint64_t l32 = tile_lanes(TILE_32x8, N), l64 = tile_lanes(TILE_64x4, N);
    return (l32 * 10 <= l64 * 9 || (l32 <= l64 && N <= 128)) ? TILE_32x8 : TILE_64x4;
}
// default planes per block of the march (64*KW x 4 tiles): N halved while the
// launch has fewer than `target` blocks, but never below 16 planes (or N) (the
// 3-plane prologue costs 3/DI extra plane reads). The storage policies of the march
// set the target (S64::MARCH_BLOCKS, S32::MARCH_BLOCKS), from the 512^3 gauged
// step timed on one grid with alternating depths: FP32C 8.43 / 8.14 / 8.01 /
// 8.07 ms at DI = 128 / 64 / 32 / 16 (4096 ... 32768 blocks), FP64 15.68 /
// 15.59 / 15.80 / 16.24 ms.
inline int march_di(int N, int KW, uint32_t target = 4096) {
    const uint32_t tiles = (uint32_t)((N + 64 * KW - 1) / (64 * KW)) * march_tiles_j(N);
    int di = N;
    while (di > 16 && (uint64_t)tiles * (uint64_t)((N + di - 1) / di) < target / (uint32_t)KW) di /= 2;
    return di < 16 ? (N < 16 ? N : 16) : di;
}
// real-mode pass 1 (positions pin, velocities vel the force sees; same layout)
int launch_real_aux(const KParams &kp, const double *pin, const double *vel, double *aux, hipStream_t s);
// gradient-pinned boundary: the y-edge rows after a K_STEP launch of the
// direct FP64 kernel (nf fields; b: that launch's buffers)
int launch_bc1_edges(bool emit, int nf, const KParams &kp, const Bufs<double> &b, hipStream_t s);
int launch_real_aux_f32(const KParams &kp, const float *pin, const float *vel, double *aux, hipStream_t s);

// Diagnostics: pass 1 writes per-block partials; the finish kernel reduces
// them in a fixed order into out (deterministic).
enum {
    DS_EPK = 0, DS_ETK, DS_EG, DS_EM, DS_EP, DS_ETG, DS_ETM, DS_EC, DS_EEM_E, DS_PINT, DS_TH2,
    DS_QP, DS_QT, DS_QC, DS_QP0, DS_QP1, DS_QP2, DS_M0, DS_CX, DS_CY, DS_CZ,
    DS_GL2, DS_NOMEGA, DS_EEM_G, DS_FSUM, DS_NC,
    DS_NSUM
};
enum { DM_PHI = 0, DM_P, DM_S, DM_GAUSS, DM_NMAX };
struct DiagParams {
    double R2core;      // qdiag_radius^2
    double Rint2;       // Gauss interior radius^2
    double Rcube;       // Q_flux cube half-width
    double G_offset;
    double BPRE;        // 1/(g^2 dx^4)
    double dV;
    int32_t gauged;     // g != 0
    int32_t has_gauge;  // gauge sector present (complex_gauge)
};
// Diagnostics fused into the emitting launch of the march (complex and gauged
// modes): every block writes DS_NSUM sums and DM_NMAX maxima of its sites to
// partials (one row of DS_NSUM + DM_NMAX per block, the pass-1 layout), and
// the gauged step writes each site's Gauss-law charge density to rho (N^3,
// unpadded) for the Gauss tail pass, which needs E at the backward neighbours
// that other blocks compute.
struct DiagArgs {
    DiagParams dp;
    double *rho;
    double *partials;
};
int launch_march_f64_diag(int mode, const KParams &kp, const Bufs<double> &b, int DI, const DiagArgs &da,
                          hipStream_t s);
int launch_march_f32c_diag(int mode, int lo_kind, const KParams &kp, const Bufs32 &b, int DI, const DiagArgs &da,
                           hipStream_t s);
// blocks of one march launch at N (64 x 4 tiles, DI planes; DI <= 0: the default)
inline int march_blocks(int N, int DI) {
    if (DI <= 0) DI = march_di(N, 1);
    return (int)(march_tiles_k(N) * march_tiles_j(N) * (uint32_t)((N + DI - 1) / DI));
}
// Gauss statistics (DS_GL2, DS_NOMEGA, DS_FSUM, DS_NC, DM_GAUSS; other slots
// zero) of the seam sites of a fused launch with segments of DI planes (<= 0:
// the default), from the emitted E and the rho kept for them, per block in the
// pass-1 layout
int launch_gauss_seam(const KParams &kp, const DiagParams &dp, int DI, const double *vo, const double *rho,
                      double *partials, int nblocks, hipStream_t s);
int launch_diag_pass1(int mode, const KParams &kp, const DiagParams &dp, const double *xo,
                      const double *vo, double *partials, int nblocks, hipStream_t s);
int launch_diag_pass2(const KParams &kp, const double *xo, double thr, double cx, double cy,
                      double cz, double *partials, unsigned long long *argmin, int nblocks,
                      hipStream_t s);
int launch_reduce(const double *partials, int nblocks, int nsum, int nmax, double *out,
                  hipStream_t s);
// the same, one block per slot (for thousands of rows); maxima of non-negative values
int launch_reduce_wide(const double *partials, int nblocks, int nsum, int nmax, double *out,
                       hipStream_t s);

// Gauss-law helpers (FP64, device CG).
int launch_gauss_residual(const KParams &kp, const double *xo, const double *vo, double *b,
                          hipStream_t s);
int launch_neg_lap(const KParams &kp, const double *p, double *ap, hipStream_t s);
int launch_add_grad(const KParams &kp, const double *chi, double *E, hipStream_t s);
int launch_dot_partials(const double *a, const double *b, uint32_t n, double *partials,
                        int nblocks, hipStream_t s);
int launch_axpy_cg(double alpha, const double *p, const double *ap, double *x, double *r,
                   uint32_t n, double *partials, int nblocks, hipStream_t s);
int launch_cg_update(double xbar, double beta, const double *r, double *x, double *p,
                     uint32_t n, hipStream_t s);
int launch_sub_mean_max(double *b, double mean, uint32_t n, double *partials, int nblocks,
                        hipStream_t s);
int launch_scale(double *x, double f, uint32_t n, hipStream_t s);

constexpr int DIAG_BLOCK = 256;

}  // namespace scp
