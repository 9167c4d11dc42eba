// scp_march.h — the LDS-staged z-march form of the fused leapfrog step
// (K_STEP launches of the complex and gauged modes, FP64 and FP32C storage).
//
// Same schedule and per-site arithmetic as scp_step.hip / scp_step32.hip (the
// expression trees of scp_sim.c:931-994, 1053-1263, 1583-1716); what changes is
// where the neighbour values come from:
//
//   * a block is a 64*KW(k) x 4(j) tile that marches along i over DI planes
//     (KW = 1 in production: two FP32C blocks or one FP64 block per CU);
//   * the own column of planes i-1, i, i+1 lives in registers (the march queue);
//   * plane i of the tile plus a one-cell halo ring (rows j0-1, j0+4, columns
//     k0-1, k0+64; of the corners, the two link diagonals the gauged forces
//     read) is staged in LDS, so the j/k neighbours and the in-plane link
//     diagonals are LDS reads instead of global loads;
//   * three LDS plane buffers rotate: while plane i is read, plane i+1 (whose
//     tile and halo are needed for the two staple diagonals th_1(i+1, j-1, k)
//     and th_2(i+1, j, k-1)) sits beside it, and the buffer of plane i-2 is
//     refilled with plane i+1 at the top of the iteration — one barrier per
//     plane;
//   * the own column and the halo of plane i+2 are loaded into registers one
//     iteration ahead, so HBM latency is covered by a whole plane of compute.
//
// Global traffic per site and step: the state read and written once plus the
// halo ring (per field, two rows of 64*KW cells and at most 2*6 column cells
// per 256*KW sites; the step launches load only the column cells their forces
// read), served by L2 where the adjacent tiles stream the same lines at about
// the same time.
//
// Ragged tiles (N not a multiple of 64 or 4) keep the scheme: the right halo
// column sits at column W+1 and the bottom halo row at row H+1, where W x H is
// the valid part of the tile, so every valid site finds its wrapped neighbour
// at the same relative LDS offset.
#pragma once

#include "scp_comp32.h"
#include "scp_diag_site.h"
#include "scp_physics.h"

#include <hip/hip_runtime.h>

namespace scp {

constexpr int MTJ = 4;                 // j rows per block (the single-grid launches)
constexpr int MNBUF = 3;

// Tile geometry: KW waves along k per row (a row of 64*KW sites; wider rows
// give the HBM longer contiguous runs per field), or KW = 0: rows of 32 sites,
// two rows per wave (grids whose edge is not a multiple of 64 waste fewer
// lanes); TJ rows (4 for single grids; batches may pick 2 or 8).
template <int KW, int TJ = MTJ>
struct MG {
    static constexpr int TKM = KW == 0 ? TK / 2 : TK * KW;   // sites along k
    static constexpr int ROWS = TJ;            // rows of the tile
    static constexpr int MR = TJ + 2;          // staged rows
    static constexpr int MC = TKM + 2;         // staged columns
    static constexpr int MPL = MR * MC;        // staged cells per field
    static constexpr int MNT = TKM * TJ;       // threads per block
    static constexpr int NW = MNT / 64;        // waves per block
};

struct MTile {
    int N, k0, j0, W, H, i0, i1, j, k, own;
    int tid;          // thread of the tile (row-major: tid = row * TKM + column)
    bool valid;
};

__device__ inline int wrapc(int x, int N) { return x < 0 ? x + N : (x >= N ? x - N : x); }

// tile (tk, tj) of segment seg of an N^3 grid; see MG for the tile size. BX:
// the launch's block width; a block of BX x (MNT / BX) threads runs the tile
// with its linear thread id (so one launch can run tiles of either 256-thread
// shape, 64 x 4 or 32 x 8)
template <int KW, int TJ = MTJ, int BX = MG<KW, TJ>::TKM>
__device__ inline MTile mtile(int N, int DI, uint32_t tk, uint32_t tj, uint32_t seg) {
    using G = MG<KW, TJ>;
    MTile t;
    int tx, ty;
    if constexpr (BX == G::TKM) {
        tx = threadIdx.x;
        ty = threadIdx.y;
    } else {
        const int lin = threadIdx.y * BX + threadIdx.x;
        tx = lin % G::TKM;
        ty = lin / G::TKM;
    }
    t.N = N;
    t.k0 = tk * G::TKM;
    t.j0 = tj * TJ;
    t.W = min(G::TKM, N - t.k0);
    t.H = min(TJ, N - t.j0);
    t.i0 = seg * DI;
    t.i1 = min(N, t.i0 + DI);
    t.k = t.k0 + tx;
    t.j = t.j0 + ty;
    t.tid = ty * G::TKM + tx;
    t.valid = tx < t.W && ty < t.H;
    t.own = (ty + 1) * G::MC + tx + 1;
    return t;
}

// Byte-offset addressing: every array is fields of fs elements; a load names a
// wave-uniform field base (scalar registers) and a 32-bit byte offset in the
// field (one vector register), so the march needs no 64-bit vector address
// arithmetic. Valid while one field stays below 4 GiB (checked on the host).
template <typename V>
__device__ inline const V *fbase(const V *p, int f, uint32_t fs) { return p + (size_t)f * fs; }
template <typename V>
__device__ inline V *fbase(V *p, int f, uint32_t fs) { return p + (size_t)f * fs; }
// Every access goes through a global-address-space pointer: the batched
// launches take their buffer pointers from a descriptor in device memory, and
// with generic pointers the compiler emits flat loads and stores, which also
// count against the LDS wait counter, so each wait for an LDS read would wait
// for the plane-ahead loads as well (the single-grid launch gets global
// accesses from its kernel arguments either way).
template <typename V>
using gptr = __attribute__((address_space(1))) V *;
template <typename V>
__device__ inline gptr<const V> gaddr(const V *base, uint32_t e) {
    return (gptr<const V>)((gptr<const char>)(base) + e * (uint32_t)sizeof(V));
}
template <typename V>
__device__ inline gptr<V> gaddr(V *base, uint32_t e) {
    return (gptr<V>)((gptr<char>)(base) + e * (uint32_t)sizeof(V));
}
template <typename V>
__device__ inline V ldo(const V *base, uint32_t e) { return *gaddr(base, e); }
template <typename V>
__device__ inline V ldo_nt(const V *base, uint32_t e) { return __builtin_nontemporal_load(gaddr(base, e)); }
template <typename V>
__device__ inline void sto_nt(V *base, uint32_t e, V v) { __builtin_nontemporal_store(v, gaddr(base, e)); }
// element idx of a global array (64-bit index)
template <typename V>
__device__ inline gptr<V> gel(V *p, size_t idx) { return (gptr<V>)(p) + idx; }
template <typename V>
__device__ inline gptr<const V> gel(const V *p, size_t idx) { return (gptr<const V>)(p) + idx; }

// The halo ring of one plane, spread over the block.
//   rows (j0-1 and j0+H): 64-site segments (field, side, half), segment
//     s = wave + NW*q, lanes along k — a wave-uniform field and a coalesced
//     row segment per load; with 32-site rows (KW = 0) one load of a wave
//     covers both rows of field s, lanes 0-31 the top row, 32-63 the bottom;
//   columns (k0-1 and k0+W): (field, side, row) cells tid, tid + MNT, ...
//     (CQ per thread), only those the step's reads need — each cell is a
//     128-byte line that the neighbouring tile streams, and the fewer lines,
//     the fewer that miss L2 and are read again from memory:
//     HALO_FACE: every field on both sides, rows j0 .. j0+H-1: the real and
//       complex forces, and the diagnostics of every mode, read the four face
//       neighbours only;
//     HALO_GAUGE: the gauged forces' reads — at k-1 every field, at k+1 every
//       field but th_2, and two corners, th_2(j0+H, k0-1) (the diagonal
//       th_2(j+1, k-1) of the last row) and th_1(j0-1, k0+W) (th_1(j-1, k+1) of
//       the first row): 118 cells of a 64 x 4 tile-plane instead of the 180 of
//       the whole ring. The gauged diagnostics' face reads (th_2 at k-1, th_0
//       and th_1 at k+1 among the links) fall within the same set.
//   Column cells that are not loaded keep stale values in the staged plane and
//   are never read.
enum { HALO_FACE = 1, HALO_GAUGE = 2 };
template <int NF, int KW, int TJ, int COLS>
struct Halo {
    using G = MG<KW, TJ>;
    static constexpr bool HALF = KW == 0;
    static constexpr int NSEG = HALF ? NF : NF * 2 * KW;   // wave loads of halo rows per plane
    static constexpr int RQ = (NSEG + G::NW - 1) / G::NW;  // of them per wave
    static_assert(COLS != HALO_GAUGE || NF == F_TH + 3, "the gauged cell set names the link fields");
    // halo column cells per plane
    static constexpr int NCC = COLS == HALO_FACE ? NF * 2 * TJ : (2 * NF - 1) * TJ + 2;
    static constexpr int CQ = (NCC + G::MNT - 1) / G::MNT; // of them per thread
    int wave, lane, W;
    uint32_t rtop, rbot;      // in-plane element offsets (padded rows) of this lane's top/bottom halo cell (half 0)
    uint32_t rtop0, rbot0;    // the same at the tile's first column
    int lrow_top, lrow_bot;   // LDS offsets (field 0, half 0) of this lane's halo row cells
    int cf[CQ];               // column cells: field, or -1
    uint32_t cg[CQ];          // column cells: in-plane element offset
    int cl[CQ];               // column cells: LDS offset
    __device__ void init(const MTile &t, uint32_t pitch) {
        const int N = t.N, tid = t.tid;
        wave = __builtin_amdgcn_readfirstlane(tid / 64);
        lane = tid & 63;
        W = t.W;
        const int sl = HALF ? (lane & 31) : lane;           // lane within its row segment
        rtop0 = (uint32_t)wrapc(t.j0 - 1, N) * pitch + t.k0;
        rbot0 = (uint32_t)wrapc(t.j0 + t.H, N) * pitch + t.k0;
        rtop = rtop0 + sl;
        rbot = rbot0 + sl;
        lrow_top = sl + 1;
        lrow_bot = (t.H + 1) * G::MC + sl + 1;
#pragma unroll
        for (int c = 0; c < CQ; c++) {
            const int cell = tid + c * G::MNT;
            cf[c] = -1; cg[c] = 0; cl[c] = 0;
            if (cell < NCC) {
                // field, side (0: k0-1, 1: k0+W) and staged row r (0: j0-1 .. H+1: j0+H)
                int f, side, r, rmax = t.H;
                constexpr int NL = NF * TJ, NRT = (COLS == HALO_GAUGE ? NF - 1 : NF) * TJ;
                if (cell < NL) {
                    side = 0; f = cell / TJ; r = 1 + cell % TJ;
                } else if (cell < NL + NRT) {
                    side = 1; f = (cell - NL) / TJ; r = 1 + (cell - NL) % TJ;   // gauged: fields before th_2
                } else if (cell == NL + NRT) {
                    side = 0; f = F_TH + 2; r = t.H + 1; rmax = t.H + 1;
                } else {
                    side = 1; f = F_TH + 1; r = 0;
                }
                if (r <= rmax) {
                    cf[c] = f;
                    const int gj = wrapc(t.j0 - 1 + r, N);
                    const int ck = side ? wrapc(t.k0 + t.W, N) : wrapc(t.k0 - 1, N);
                    cg[c] = (uint32_t)gj * pitch + ck;
                    cl[c] = f * G::MPL + r * G::MC + (side ? t.W + 1 : 0);
                }
            }
        }
    }
    // the load of segment q for every lane: the segment's field and this lane's
    // offset, in bounds also where the lane or the whole segment is idle (field 0,
    // the column clamped to the tile)
    __device__ void seg_any(int q, int &f, uint32_t &g) const {
        const int sg = wave + G::NW * q;
        if constexpr (HALF) {
            f = sg < NSEG ? sg : 0;
            g = ((lane >> 5) ? rbot0 : rtop0) + (uint32_t)min(lane & 31, W - 1);
        } else {
            const int rem = sg % (2 * KW), side = rem / KW, half = rem % KW;
            f = sg < NSEG ? sg / (2 * KW) : 0;
            g = (side ? rbot0 : rtop0) + (uint32_t)min(half * 64 + lane, W - 1);
        }
    }
    // segment q of this wave: field, global offset, LDS offset; false if idle
    __device__ bool seg(int q, int &f, uint32_t &g, int &l) const {
        const int sg = wave + G::NW * q;
        if (sg >= NSEG) return false;
        if constexpr (HALF) {
            const int side = lane >> 5;
            if ((lane & 31) >= W) return false;
            f = sg;
            g = side ? rbot : rtop;
            l = side ? lrow_bot : lrow_top;
            return true;
        } else {
            f = sg / (2 * KW);
            const int rem = sg % (2 * KW), side = rem / KW, half = rem % KW;
            if (half * 64 + lane >= W) return false;
            g = (side ? rbot : rtop) + half * 64;
            l = (side ? lrow_bot : lrow_top) + half * 64;
            return true;
        }
    }
};

// own column of a staged plane, indexable by field
template <typename R, int KW, int TJ = MTJ>
struct LdsCol {
    const R *p;
    __device__ R operator[](int f) const { return p[f * MG<KW, TJ>::MPL]; }
};

// neighbour provider over the staged planes (see GlobalNB for the contract)
template <typename R, int KW, int TJ = MTJ>
struct LdsNB {
    static constexpr int MPL = MG<KW, TJ>::MPL, MC = MG<KW, TJ>::MC;
    const R *c;   // plane i buffer, at the own cell
    const R *n;   // plane i+1 buffer, at the own cell
    __device__ R face(int f, int d, int sg) const { return c[f * MPL + (d == 1 ? sg * MC : sg)]; }
    __device__ R th1_jm_kp() const { return c[(F_TH + 1) * MPL - MC + 1]; }
    __device__ R th2_jp_km() const { return c[(F_TH + 2) * MPL + MC - 1]; }
    __device__ R th1_ip_jm() const { return n[(F_TH + 1) * MPL - MC]; }
    __device__ R th2_ip_km() const { return n[(F_TH + 2) * MPL - 1]; }
};

// ---------------------------------------------------------------------------
// leapfrog updates of one scalar (identical to the direct-neighbour kernels)
// ---------------------------------------------------------------------------

// FP32C: DF, comp_add, leap32 (scp_comp32.h)

struct S64 {
    static constexpr int MINB = 1;             // the staged planes fill the LDS: one block per CU
    static constexpr uint32_t MARCH_BLOCKS = 4096;   // single-grid launch: blocks the default depth aims at
    // tile rows per strip of the single-grid tile order (march_tile_of): row
    // order; strips of four ran 3 % slower at 512^3 (one block per CU)
    static constexpr uint32_t TILE_STRIP = 1;
    // own-site words loaded with their plane: one plane ahead (30 more VGPRs) measured
    // 1-6 % slower at 512^3 with one 4-wave block per CU
    static constexpr bool WAHEAD = false;
    static constexpr bool REGQ = true;         // registers to spare: own column of planes i, i+1 in registers
    static constexpr bool SPLIT_DAMP = false;
    using T = double;
    using B = Bufs<double>;
    struct Own { double w; };
    struct Lo {};                              // no compensation words
    __device__ static const double *pin(const B &b) { return b.pin; }
    template <int NF>
    __device__ static void lo_load(const B &, uint32_t, uint32_t, Lo &) {}
    template <int NF>
    __device__ static void lo_store(const B &, uint32_t, uint32_t, const Lo &) {}
    __device__ static Own own_load(const B &b, int f, uint32_t fh, uint32_t e, const Lo &) {
        return {ldo_nt(fbase(b.w, f, fh), e)};
    }
    // fh, e: field stride and index of the step state; fs, ec: of the emitted state;
    // xe, ve: the emitted position and velocity (x_n, v_n in FP64)
    template <int MODE, bool EMIT>
    __device__ static void update(const KParams &kp, const B &b, int f, uint32_t fh, uint32_t e, uint32_t fs,
                                  uint32_t ec, double a, double x, Own s, Lo &, bool damp, double dmp,
                                  double &xe, double &ve) {
        double v = s.w;
        v += kp.hdt * a;                               // stage 4 (C:1680-1696)
        if (damp) v *= dmp;                            // stage 5 (C:1454-1460)
        const double w = v + kp.hdt * a;               // next stage 1 (C:1597-1613)
        double xn;
        if (MODE == 2 && f >= F_TH) {                  // next stage 2, links (C:1641-1651)
            const double t = x + kp.gad * w;
            xn = t - 2.0 * SCP_PI * rint(t / (2.0 * SCP_PI));
        } else {
            xn = x + kp.dt * w;                        // next stage 2 (C:1629-1638)
        }
        sto_nt(fbase(b.w, f, fh), e, w);
        sto_nt(fbase(b.pout, f, fh), e, xn);
        xe = x;
        ve = v;
        if constexpr (EMIT) {
            __builtin_nontemporal_store(x, gel(b.xo, (size_t)f * fs + ec));
            __builtin_nontemporal_store(v, gel(b.vo, (size_t)f * fs + ec));
        }
    }
};

// FP64 policy of the fused-diagnostics launch: the own column is read back
// from the staged planes instead of held in registers (same values, so the
// same results), which leaves the registers to the diagnostics.
struct S64D : S64 {
    static constexpr bool REGQ = false;
    static constexpr bool SPLIT_DAMP = true;
};

// Compensation words in the march's addressing, one word per field (LO_F32,
// LO_BF16): raw word loads (decoded against their hi word) and encoded stores.
// LO_Q8 words go through S32's packed per-site words instead.
template <int LOK> struct LoIO {
    using LW = LoWord<LOK>;
    using W = typename LW::W;
    __device__ static W raw(const void *p, int f, uint32_t fs, uint32_t e) {
        if constexpr (LW::STORED) return ldo_nt(fbase((const W *)p, f, fs), e);
        else return W{};
    }
    __device__ static void put(void *p, int f, uint32_t fs, uint32_t e, float hi, float lo) {
        if constexpr (LW::STORED) sto_nt(fbase((W *)p, f, fs), e, LW::enc(lo, hi));
    }
};

template <int LOK>
struct S32 {
    static constexpr int MINB = 2;             // two blocks per CU: registers capped at 256
    // single-grid launch: blocks the default depth aims at (512^3: 32 planes per
    // block, 5 % faster than 128; shorter marches keep adjacent tiles closer in step)
    static constexpr uint32_t MARCH_BLOCKS = 16384;
    // strips of four tile rows in the single-grid tile order (march_tile_of):
    // 512^3 step -0.6 %, 0.6 GB fewer reads per launch
    static constexpr uint32_t TILE_STRIP = 4;
    static constexpr bool WAHEAD = true;       // own-site words loaded one plane ahead (+3 % at 512^3)
    static constexpr bool REGQ = false;        // own column of planes i, i+1 read back from LDS
    static constexpr bool SPLIT_DAMP = true;
    using T = float;
    using B = Bufs32;
    static constexpr int LKW = lo_kind_w(LOK), LKP = lo_kind_p(LOK);   // word storage of W and of P
    using LWW = LoWord<LKW>;
    using LWP = LoWord<LKP>;
    static constexpr bool PACKW = LKW == LO_Q8, PACKP = LKP == LO_Q8;
    struct Own { float wh; typename LWW::W wr; typename LWP::W xr; };   // W hi word; raw lo words of W and P
    // packed compensation words of the site: W's and P's (loaded), W's and P's out (stored)
    struct Lo { LoQuads w, x; };
    __device__ static const float *pin(const B &b) { return b.pin; }
    template <int NF>
    __device__ static void lo_load(const B &b, uint32_t fh, uint32_t e, Lo &lo) {
#pragma unroll
        for (int q = 0; q < (NF + 3) / 4; q++) {
            if constexpr (PACKW) lo.w.q[q] = ldo_nt(fbase((const uint32_t *)b.wlo, q, fh), e);
            if constexpr (PACKP) lo.x.q[q] = ldo_nt(fbase((const uint32_t *)b.pinlo, q, fh), e);
        }
    }
    template <int NF>
    __device__ static void lo_store(const B &b, uint32_t fh, uint32_t e, const Lo &lo) {
#pragma unroll
        for (int q = 0; q < (NF + 3) / 4; q++) {
            if constexpr (PACKW) sto_nt(fbase((uint32_t *)b.wlo, q, fh), e, lo.w.q[q]);
            if constexpr (PACKP) sto_nt(fbase((uint32_t *)b.poutlo, q, fh), e, lo.x.q[q]);
        }
    }
    __device__ static Own own_load(const B &b, int f, uint32_t fh, uint32_t e, const Lo &lo) {
        Own o;
        o.wh = ldo_nt(fbase(b.w, f, fh), e);
        if constexpr (PACKW) o.wr = lo.w.get(f);
        else o.wr = LoIO<LKW>::raw(b.wlo, f, fh, e);
        if constexpr (PACKP) o.xr = lo.x.get(f);
        else o.xr = LoIO<LKP>::raw(b.pinlo, f, fh, e);
        return o;
    }
    // fh, e: field stride and index of the step state; fs, ec: of the emitted state;
    // xe, ve: the emitted position and velocity (hi + lo in FP64)
    template <int MODE, bool EMIT>
    __device__ static void update(const KParams &kp, const B &b, int f, uint32_t fh, uint32_t e, uint32_t fs,
                                  uint32_t ec, float a, float xh, Own s, Lo &lo, bool damp, double dmpd,
                                  double &xe, double &ve) {
        const DF hdt = df_split(kp.hdt), dt = df_split(kp.dt), gad = df_split(kp.gad);
        float wh = s.wh, wl = LWW::dec(s.wr, s.wh), xl = LWP::dec(s.xr, xh), vh, vl;
        const float xh0 = xh, xl0 = xl;
        leap32(MODE == 2 && f >= F_TH, a, damp, (float)dmpd, hdt, dt, gad, wh, wl, xh, xl, vh, vl);
        sto_nt(fbase(b.w, f, fh), e, wh);
        sto_nt(fbase(b.pout, f, fh), e, xh);
        // packed words are stored by lo_store once every field is done
        if constexpr (PACKW) lo.w.set(f, LWW::enc(wl, wh));
        else LoIO<LKW>::put(b.wlo, f, fh, e, wh, wl);
        if constexpr (PACKP) lo.x.set(f, LWP::enc(xl, xh));
        else LoIO<LKP>::put(b.poutlo, f, fh, e, xh, xl);
        xe = (double)xh0 + (double)xl0;
        ve = (double)vh + (double)vl;
        if constexpr (EMIT) {
            __builtin_nontemporal_store(xe, gel(b.xo, (size_t)f * fs + ec));
            __builtin_nontemporal_store(ve, gel(b.vo, (size_t)f * fs + ec));
        }
    }
};

// ---------------------------------------------------------------------------
// the kernel
// ---------------------------------------------------------------------------

template <int MODE>
constexpr int march_nf() { return MODE == 2 ? 15 : (MODE == 1 ? 12 : 6); }   // staged fields (matter + links)
// MODE: 0 real, 1 complex, 2 gauged, 3 real with the pass-1 fields (alpha_cs,
// beta_h, sigma_freq: an auxiliary array read from global memory; a separate
// instance so that the plain real step keeps its registers)
template <class S, int KW, int MODE, int TJ = MTJ>
constexpr int march_lds_elems() { return MNBUF * march_nf<MODE>() * MG<KW, TJ>::MPL; }

// Per-thread diagnostics accumulators of the DIAG instance (DS_* sums, DM_* maxima).
struct DiagState {
    double S[DS_NSUM], M[DM_NMAX];
    double *ex;   // gauged: LDS row exchange of E_1 between the tile's waves (2 x MTJ x 64 doubles)
};

// The Gauss residual of a fused gauged launch, one site per thread, completed
// one plane late: G = div E - g rho needs E_d(x - e_d). E_0(x - e0) is carried
// from plane i-1 of the column, E_2(x - e2) comes from the next lower lane, and
// E_1(x - e1) from the row below through LDS, written by its wave in the
// plane's iteration and read after the next iteration's barrier (two
// alternating rows of slots, no barrier of its own). Sites whose backward
// neighbour another block emits (the segment's first plane, the tile's first
// row and column) keep rho for the seam pass (k_gauss_seam), which forms G
// from the emitted state in the same expression.
struct GaussPending {
    double E[3], E0m, E2m, rho;
    uint32_t ec;
    int i;
};

__device__ inline void gauss_pending(const KParams &kp, const DiagArgs *da, DiagState *ds, const MTile &t,
                                     const GaussPending &pg) {
    if (!t.valid) return;
    const int row = threadIdx.y, lane = threadIdx.x;
    if (pg.i > t.i0 && row > 0 && lane > 0) {
        const double Em[3] = {pg.E0m, ds->ex[((pg.i & 1) * MTJ + row - 1) * 64 + lane], pg.E2m};
        gauss_site(kp, da->dp, pg.E, Em, pg.rho, pg.i, t.j, t.k, ds->S, ds->M);
    } else {
        __builtin_nontemporal_store(pg.rho, gel(da->rho, pg.ec));
    }
}

// One tile of one grid: the march over planes [seg*DI, seg*DI + DI) of the
// 64*KW x 4 column tile (tk, tj). lds: the block's staging area,
// march_lds_elems() words. Ends with a barrier, so the caller may stage the
// next tile into the same area. DIAG (with EMIT): the emitted state's per-site
// diagnostics are accumulated into ds, and the gauged charge density is
// written to da->rho.
template <class S, int KW, int MODE, bool EMIT, bool DIAG = false, int TJ = MTJ, int BX = MG<KW, TJ>::TKM>
__device__ __forceinline__ void march_tile(const KParams &kp, const typename S::B &b, int DI, uint32_t tk,
                                           uint32_t tj, uint32_t seg, typename S::T *lds,
                                           const DiagArgs *da = nullptr, DiagState *ds = nullptr) {
    using T = typename S::T;
    constexpr int NF = march_nf<MODE>();
    constexpr int NM = MODE == 2 ? 12 : NF;   // fields of the i-1 plane (gauged: matter only)
    constexpr int MPL = MG<KW, TJ>::MPL;
    // the halo column cells the launch's reads need
    using HS = Halo<NF, KW, TJ, MODE == 2 ? HALO_GAUGE : HALO_FACE>;
    constexpr int RQ = HS::RQ, CQ = HS::CQ;
    static_assert(!DIAG || (KW == 1 && TJ == MTJ && BX == 64), "the fused diagnostics run on 64 x 4 blocks");

    const int N = kp.N;
    const MTile t = mtile<KW, TJ, BX>(N, DI, tk, tj, seg);
    if (t.i0 >= N) return;                    // whole block: uniform
    HS hs;
    hs.init(t, kp.pitch);
    // step state: padded rows (kp.pitch), plane stride kp.plane, field stride kp.fsh;
    // emitted state: unpadded (N, N*N, kp.fs)
    uint32_t fs = kp.fsh;              // made opaque per plane in the loop below
    const uint32_t NN = kp.plane;
    const T *__restrict__ pin = S::pin(b);
    const uint32_t col = t.valid ? (uint32_t)t.j * kp.pitch + (uint32_t)t.k : 0u;
    const uint32_t colc = t.valid ? (uint32_t)t.j * (uint32_t)N + (uint32_t)t.k : 0u;
    auto ld = [&](int f, uint32_t e) -> T { return ldo(fbase(pin, f, fs), e); };
    // halo loads of one plane: row pairs (wave-uniform field) and this thread's column cell
    auto halo_rows = [&](uint32_t pe, T *hr) {
#pragma unroll
        for (int q = 0; q < RQ; q++) {
            int f, l;
            uint32_t g;
            hr[q] = hs.seg(q, f, g, l) ? ld(f, pe + g) : T(0);
        }
    };
    auto halo_col = [&](uint32_t pe, T *hc) {
#pragma unroll
        for (int c = 0; c < CQ; c++) hc[c] = hs.cf[c] >= 0 ? *gel(pin, (size_t)hs.cf[c] * fs + pe + hs.cg[c]) : T(0);
    };
    // the same, loading on every lane and segment (idle ones from in-bounds places
    // whose values publish does not stage)
    auto halo_rows_any = [&](uint32_t pe, T *hr) {
#pragma unroll
        for (int q = 0; q < RQ; q++) {
            int f;
            uint32_t g;
            hs.seg_any(q, f, g);
            hr[q] = ld(f, pe + g);
        }
    };
    auto halo_col_any = [&](uint32_t pe, T *hc) {
#pragma unroll
        for (int c = 0; c < CQ; c++)
            hc[c] = *gel(pin, (size_t)(hs.cf[c] < 0 ? 0 : hs.cf[c]) * fs + pe + hs.cg[c]);
    };
    auto publish = [&](T *buf, const T *own, const T *hr, const T *hc) {
        if (t.valid) {
#pragma unroll
            for (int f = 0; f < NF; f++) buf[f * MPL + t.own] = own[f];
        }
#pragma unroll
        for (int q = 0; q < RQ; q++) {
            int f, l;
            uint32_t g;
            if (hs.seg(q, f, g, l)) buf[f * MPL + l] = hr[q];
        }
#pragma unroll
        for (int c = 0; c < CQ; c++)
            if (hs.cf[c] >= 0) buf[hs.cl[c]] = hc[c];
    };

    // prologue: planes i0-1 (own column), i0 (own + halo), i0+1 (own + halo)
    const int im0 = wm(t.i0, N), ip0 = wp(t.i0, N);
    T cm[NM], cp[NF], hr[RQ], hc[CQ];
    T cq[S::REGQ ? NF : 1];     // register copy of plane i (register-queue policy)
    GaugeCarry<T> cy{};
    T th0m = T(0);              // DIAG, gauged: th_0 at plane i-1 (backward transport along i)
    // DIAG, gauged: this site's emitted E and charge density (the Gauss residual is
    // completed after the compute, with the backward neighbours' E), and E_0 of the
    // own column at plane i-1
    double gE[3] = {0.0, 0.0, 0.0}, grho = 0.0;
    GaussPending pg{};          // DIAG, gauged: the previous plane's site
    {
        const uint32_t em = (uint32_t)im0 * NN + col, e0 = (uint32_t)t.i0 * NN + col, ep = (uint32_t)ip0 * NN + col;
        T c0[NF];
        if (t.valid) {
#pragma unroll
            for (int f = 0; f < NM; f++) cm[f] = ld(f, em);
#pragma unroll
            for (int f = 0; f < NF; f++) { c0[f] = ld(f, e0); cp[f] = ld(f, ep); }
            if constexpr (DIAG && MODE == 2) th0m = ld(F_TH + 0, em);
            if constexpr (MODE == 2) {
                const uint32_t jpo = (uint32_t)im0 * NN + (uint32_t)wp(t.j, N) * kp.pitch + t.k;
                const uint32_t kpo = (uint32_t)im0 * NN + (uint32_t)t.j * kp.pitch + wp(t.k, N);
                gauge_prologue(ld(F_TH + 0, em), ld(F_TH + 1, em), ld(F_TH + 2, em), c0[F_TH + 1], c0[F_TH + 2],
                               ld(F_TH + 0, jpo), ld(F_TH + 0, kpo), cy);
            }
        } else {
#pragma unroll
            for (int f = 0; f < NM; f++) cm[f] = T(0);
#pragma unroll
            for (int f = 0; f < NF; f++) { c0[f] = T(0); cp[f] = T(0); }
        }
        T h0r[RQ], h0c[CQ];
        halo_rows((uint32_t)t.i0 * NN, h0r);
        halo_col((uint32_t)t.i0 * NN, h0c);
        halo_rows((uint32_t)ip0 * NN, hr);
        halo_col((uint32_t)ip0 * NN, hc);
        publish(lds, c0, h0r, h0c);
        if constexpr (S::REGQ) {
#pragma unroll
            for (int f = 0; f < NF; f++) cq[f] = c0[f];
        }
    }

    // own-site words (W and compensation words) of the plane at element offset e
    auto load_own = [&](uint32_t e, typename S::Own *o) {
        typename S::Lo li;
        S::template lo_load<NF>(b, fs, e, li);
#pragma unroll
        for (int f = 0; f < NF; f++) o[f] = S::own_load(b, f, fs, e, li);
    };
    // INPLACE (own column read back from LDS, own-site words one plane ahead):
    // the prefetches land in the registers their consumer reads — plane i+2 into
    // cp / hr / hc right after cp / hr / hc of plane i+1 are staged, and the
    // own-site words of plane i+1 into sw after plane i's update — so no
    // loop-carried copy of a load result (which makes the compiler wait at the
    // loop end for every memory operation in flight, stores included) is left.
    // Those loads are issued on every lane and plane (idle ones from in-bounds
    // places), so no branch around them leaves the compiler unsure how many
    // memory operations are in flight: it then waits for the staged plane's
    // loads alone, not for everything. FP32C 512^3 step: 7.35 ms against 7.55
    // (the loads behind lane branches) and 7.75-7.99 (copies).
    constexpr bool INPLACE = !S::REGQ && S::WAHEAD;
    typename S::Own swn[S::WAHEAD && !INPLACE ? NF : 1];
    typename S::Own swi[INPLACE ? NF : 1];
    if constexpr (INPLACE) {
        if (t.valid && t.i0 < t.i1) load_own((uint32_t)t.i0 * NN + col, swi);
    } else if constexpr (S::WAHEAD) {
        if (t.valid && t.i0 < t.i1) load_own((uint32_t)t.i0 * NN + col, swn);
    }

    int bb = 0;
    for (int i = t.i0; i < t.i1; i++) {
        // The field stride, opaque per plane (an empty asm that "writes" its SGPR):
        // the 45-60 field bases are then formed where used, by scalar adds, instead
        // of being hoisted out of the loop and spilled to VGPR lanes (-4.5 % FP32C,
        // -1 to -4 % FP64 step time at 512^3).
        asm volatile("" : "+s"(fs));
        const int nbuf = (bb + 1 == MNBUF) ? 0 : bb + 1;
        // plane i+2; on the segment's last plane (whose prefetch is never staged) plane
        // i+1 again, loaded the plane before, so the unused loads read lines L2 still holds
        const uint32_t p2e = (uint32_t)(i + 1 < t.i1 ? wp(wp(i, N), N) : wp(i, N)) * NN;
        const uint32_t e = (uint32_t)i * NN + col;
        const uint32_t ec = (uint32_t)i * (uint32_t)N * (uint32_t)N + colc;
        // prefetch plane i+2 (own column + halo) and the own-site velocity words of
        // plane i+1 (WAHEAD: a whole plane of compute to arrive) or of plane i
        T cpp[INPLACE ? 1 : NF], hrr[INPLACE ? 1 : RQ], hcc[INPLACE ? 1 : CQ];
        typename S::Own swc[INPLACE ? 1 : NF];
        typename S::Own *sw = INPLACE ? swi : swc;
        if constexpr (INPLACE) {
            // plane i+1 into the buffer that held plane i-2, then plane i+2 into its registers
            publish(lds + nbuf * NF * MPL, cp, hr, hc);
            // on every lane (col is 0, an in-bounds site, where the lane is idle)
#pragma unroll
            for (int f = 0; f < NF; f++) cp[f] = ld(f, p2e + col);
            halo_rows_any(p2e, hr);
            halo_col_any(p2e, hc);
        } else {
            if constexpr (S::WAHEAD) {
#pragma unroll
                for (int f = 0; f < NF; f++) swc[f] = swn[f];
            }
            if (t.valid) {
#pragma unroll
                for (int f = 0; f < NF; f++) cpp[f] = ld(f, p2e + col);
                if constexpr (S::WAHEAD) {
                    if (i + 1 < t.i1) load_own(e + NN, swn);
                } else {
                    load_own(e, swc);
                }
            } else {
#pragma unroll
                for (int f = 0; f < NF; f++) cpp[f] = T(0);
            }
            halo_rows(p2e, hrr);
            halo_col(p2e, hcc);
            // plane i+1 into the buffer that held plane i-2
            publish(lds + nbuf * NF * MPL, cp, hr, hc);
        }
        __syncthreads();
        if constexpr (DIAG && MODE == 2) {
            if (i > t.i0) gauss_pending(kp, da, ds, t, pg);
        }

        // own column of planes i and i+1: registers (REGQ) or read back from the staged tiles
        auto compute = [&](const auto &c0, const auto &c1) {
            const LdsNB<T, KW, TJ> nb{lds + bb * NF * MPL + t.own, lds + nbuf * NF * MPL + t.own};
            T acc[NF];
            if constexpr (MODE == 2)
                forces_gauge(kp, c0, c1, cm, nb, cy, acc);
            else if constexpr (MODE == 1)
                forces_complex(kp, c0, c1, cm, nb, acc);
            else if constexpr (MODE == 0)
                forces_real(kp, c0, c1, cm, nb, [](int, uint32_t) { return T(0); },
                            [](int, int) { return 0u; }, acc);
            else {
                // real mode with the pass-1 fields: they come from global memory at the unpadded
                // site index, which equals the step-state index here (real-mode rows are not
                // padded: pitch = N)
                const double *__restrict__ aux = b.aux;
                const auto AUX = [&](int f, uint32_t idx) -> T { return T(*gel(aux, (size_t)f * kp.fs + idx)); };
                // neighbour indices formed where a pass-1 term reads them (not held otherwise)
                const auto nbi = [&](int d, int sg) -> uint32_t {
                    if (d == 3) return e;
                    if (d == 0) return (uint32_t)(sg > 0 ? wp(i, N) : wm(i, N)) * NN + col;
                    const uint32_t pl = (uint32_t)i * NN;
                    if (d == 1) return pl + (uint32_t)(sg > 0 ? wp(t.j, N) : wm(t.j, N)) * kp.pitch + t.k;
                    return pl + (uint32_t)t.j * kp.pitch + (sg > 0 ? wp(t.k, N) : wm(t.k, N));
                };
                forces_real(kp, c0, c1, cm, nb, AUX, nbi, acc);
            }
            double dmp = 1.0;
            double xe[NF], ve[NF];   // the emitted state of this site (used by DIAG)
            typename S::Lo lo{};     // packed compensation words out
            if constexpr (S::SPLIT_DAMP) {
                // separate damped / undamped paths: a select between them would keep both live
                if (kp.damp && sponge(kp, i, t.j, t.k, dmp)) {
#pragma unroll
                    for (int f = 0; f < NF; f++)
                        S::template update<MODE, EMIT>(kp, b, f, fs, e, kp.fs, ec, acc[f], c0[f], sw[f], lo, true, dmp,
                                                       xe[f], ve[f]);
                } else {
#pragma unroll
                    for (int f = 0; f < NF; f++)
                        S::template update<MODE, EMIT>(kp, b, f, fs, e, kp.fs, ec, acc[f], c0[f], sw[f], lo, false, 1.0,
                                                       xe[f], ve[f]);
                }
            } else {
                const bool damp = kp.damp && sponge(kp, i, t.j, t.k, dmp);
#pragma unroll
                for (int f = 0; f < NF; f++)
                    S::template update<MODE, EMIT>(kp, b, f, fs, e, kp.fs, ec, acc[f], c0[f], sw[f], lo, damp, dmp,
                                                   xe[f], ve[f]);
            }
            S::template lo_store<NF>(b, fs, e, lo);
            if constexpr (DIAG) {
                // neighbours at the stencil level of this launch: plane i+1 (own column),
                // plane i-1 (matter column and th_0), in-plane from the staged tile
                const auto xn = [&](int f, int d, int sg) -> double {
                    if (d == 0) {
                        if (sg > 0) return (double)c1[f];
                        return f < NM ? (double)cm[f] : (double)th0m;
                    }
                    return (double)nb.face(f, d, sg);
                };
                const double rho = diag_site<MODE == 3 ? 0 : MODE>(kp, da->dp, xe, ve, xn, i, t.j, t.k, ds->S, ds->M);
                if constexpr (MODE == 2) {
                    gE[0] = ve[F_TH]; gE[1] = ve[F_TH + 1]; gE[2] = ve[F_TH + 2];
                    grho = rho;
                }
            }
        };
        if constexpr (S::REGQ) {
            if (t.valid) compute(cq, cp);
#pragma unroll
            for (int f = 0; f < NM; f++) cm[f] = cq[f];
            if constexpr (DIAG && MODE == 2) th0m = cq[F_TH];
#pragma unroll
            for (int f = 0; f < NF; f++) { cq[f] = cp[f]; cp[f] = cpp[f]; }
        } else {
            const LdsCol<T, KW, TJ> c0{lds + bb * NF * MPL + t.own}, c1{lds + nbuf * NF * MPL + t.own};
            if (t.valid) compute(c0, c1);
#pragma unroll
            for (int f = 0; f < NM; f++) cm[f] = c0[f];
            if constexpr (DIAG && MODE == 2) th0m = c0[F_TH];
            if constexpr (INPLACE) {
                // on every lane and plane (the segment's last reloads its own plane's words, unused)
                load_own((uint32_t)(i + 1 < t.i1 ? wp(i, N) : i) * NN + col, swi);
            } else {
#pragma unroll
                for (int f = 0; f < NF; f++) cp[f] = cpp[f];
            }
        }
        if constexpr (!INPLACE) {
#pragma unroll
            for (int q = 0; q < RQ; q++) hr[q] = hrr[q];
#pragma unroll
            for (int c = 0; c < CQ; c++) hc[c] = hcc[c];
        }
        bb = nbuf;
        if constexpr (DIAG && MODE == 2) {
            // this plane's Gauss site waits for the row below's E_1 (LDS, visible after
            // the next barrier); E_2 of the next lower lane comes by shuffle now
            const double e2m = __shfl_up(gE[2], 1, 64);
            if (t.valid) ds->ex[((i & 1) * MTJ + threadIdx.y) * 64 + threadIdx.x] = gE[1];
            pg.E0m = pg.E[0];
            pg.E[0] = gE[0]; pg.E[1] = gE[1]; pg.E[2] = gE[2];
            pg.rho = grho;
            pg.E2m = e2m;
            pg.ec = ec;
            pg.i = i;
        }
    }
    __syncthreads();
    if constexpr (DIAG && MODE == 2) {
        if (t.i0 < t.i1) gauss_pending(kp, da, ds, t, pg);
    }
}

// Tile (tk, tj) and segment (returned) of logical block lb. Within a segment the
// tiles go in strips of SH j-adjacent tile rows, j fastest in a strip: a tile's
// j neighbours in its strip come one block before or after it and its k
// neighbours SH, so with the XCD-aware numbering (an XCD's blocks take
// consecutive logical blocks) more of the halo rows a tile loads are lines its
// neighbours stream into the same L2 at about the same time (with row order,
// SH = 1, the j neighbours are a whole tile row of blocks apart). Row order when
// the tile rows do not split into strips of SH.
template <uint32_t SH>
__device__ inline uint32_t march_tile_of(uint32_t lb, uint32_t ntk, uint32_t ntj, uint32_t &tk, uint32_t &tj) {
    const uint32_t t = lb % (ntk * ntj);
    if (ntj % SH == 0) {
        const uint32_t r = t % (SH * ntk);
        tk = r / SH;
        tj = t / (SH * ntk) * SH + r % SH;
    } else {
        tk = t % ntk;
        tj = t / ntk;
    }
    return lb / (ntk * ntj);
}

// One grid: block b runs tile (tk, tj, seg) of the XCD-aware logical numbering.
template <class S, int KW, int MODE, bool EMIT>
__global__ void __launch_bounds__(MG<KW>::MNT, KW == 1 ? S::MINB : 1)
k_march(KParams kp, typename S::B b, int DI) {
    __shared__ typename S::T lds[march_lds_elems<S, KW, MODE>()];
    const uint32_t ntk = (kp.N + MG<KW>::TKM - 1) / MG<KW>::TKM, ntj = (kp.N + MTJ - 1) / MTJ;
    uint32_t tk, tj;
    const uint32_t lb = march_tile_of<S::TILE_STRIP>(logical_block(blockIdx.x, gridDim.x), ntk, ntj, tk, tj);
    march_tile<S, KW, MODE, EMIT, false, MTJ, MG<KW>::TKM>(kp, b, DI, tk, tj, lb, lds);
}

// The emitting launch with the diagnostics fused (single grid, KW = 1): each
// block reduces its threads' accumulators (wave butterfly, then its waves in
// order: deterministic) into its row of da.partials, also when it has no
// planes (the row is then zero).
// One block per CU: the diagnostics need the registers (the FP32C instance
// capped for two blocks per CU spills and ran 1.65x slower at 512^3).
template <class S, int MODE>
__global__ void __launch_bounds__(MG<1>::MNT, 1)
k_march_diag(KParams kp, typename S::B b, int DI, DiagArgs da) {
    __shared__ typename S::T lds[march_lds_elems<S, 1, MODE>()];
    constexpr int NQ = DS_NSUM + DM_NMAX, NW = MG<1>::NW;
    const uint32_t ntk = (kp.N + MG<1>::TKM - 1) / MG<1>::TKM, ntj = (kp.N + MTJ - 1) / MTJ;
    uint32_t tk, tj;
    const uint32_t lb = march_tile_of<S::TILE_STRIP>(logical_block(blockIdx.x, gridDim.x), ntk, ntj, tk, tj);
    __shared__ double ex[MODE == 2 ? 2 * MTJ * 64 : 1];
    DiagState ds;
    ds.ex = ex;
#pragma unroll
    for (int q = 0; q < DS_NSUM; q++) ds.S[q] = 0.0;
#pragma unroll
    for (int q = 0; q < DM_NMAX; q++) ds.M[q] = 0.0;
    march_tile<S, 1, MODE, true, true, MTJ, MG<1>::TKM>(kp, b, DI, tk, tj, lb, lds, &da, &ds);
    // march_tile ended with a barrier, or the whole block returned before staging: lds is free
    double *red = reinterpret_cast<double *>(lds);
    const int tid = threadIdx.y * MG<1>::TKM + threadIdx.x, lane = tid & 63, wv = tid >> 6;
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        double v = q < DS_NSUM ? ds.S[q] : ds.M[q - DS_NSUM];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const double y = __shfl_xor(v, o, 64);
            v = q < DS_NSUM ? v + y : fmax(v, y);
        }
        if (lane == 0) red[wv * NQ + q] = v;
    }
    __syncthreads();
    if (tid < NQ) {
        double v = red[tid];
        for (int w = 1; w < NW; w++) v = tid < DS_NSUM ? v + red[w * NQ + tid] : fmax(v, red[w * NQ + tid]);
        da.partials[(size_t)blockIdx.x * NQ + tid] = v;
    }
}

template <class S, int MODE>
inline int launch_march_diag_t(const KParams &kp, const typename S::B &b, int DI, const DiagArgs &da,
                               hipStream_t s) {
    if (DI <= 0) DI = march_di(kp.N, 1);
    const dim3 grid((uint32_t)march_blocks(kp.N, DI)), block(MG<1>::TKM, MTJ);
    hipLaunchKernelGGL((k_march_diag<S, MODE>), grid, block, 0, s, kp, b, DI, da);
    return (int)hipGetLastError();
}

// Many grids: block b runs tiles [b*per_block, (b+1)*per_block) of the table
// (b in XCD-aware order, so an XCD's blocks take adjacent tiles); a grid takes
// part in chain launches q < nq and emits its FP64 state on the last (if it
// holds the emit buffers: grids of the request loop release them in between).
// Tiles are 64*KW x TJ sites (KW = 0: 32 x TJ); blocks per CU: the default
// 64 x 4 tile and the 32 x 8 tile (61 KB of FP32 planes) fit S::MINB, 64 x 2
// three, the larger ones one.
template <class S, int KW, int TJ>
constexpr int march_minb() {
    return ((KW == 1 && TJ == MTJ) || (KW == 0 && TJ == 8)) ? S::MINB : ((KW == 1 && TJ < MTJ) ? 3 : 1);
}

template <class S, int KW, int MODE, int TJ = MTJ>
__global__ void __launch_bounds__((MG<KW, TJ>::MNT), (march_minb<S, KW, TJ>()))
k_march_batch(const BatchGrid<typename S::B> *__restrict__ grids, const BatchTile *__restrict__ tiles,
              uint32_t ntiles, uint32_t per_block, int q) {
    __shared__ typename S::T lds[march_lds_elems<S, KW, MODE, TJ>()];
    const uint32_t b = logical_block(blockIdx.x, gridDim.x);
    const uint32_t t1 = min(ntiles, (b + 1) * per_block);
    for (uint32_t t = b * per_block; t < t1; t++) {
        const BatchTile bt = tiles[t];
        const BatchGrid<typename S::B> &g = grids[bt.grid];
        if (q >= g.nq) continue;
        const uint32_t tk = bt.tile % g.ntk, r = bt.tile / g.ntk;
        const uint32_t tj = r % g.ntj, seg = r / g.ntj;
        const int par = (g.cur0 + q) & 1;
        // the grid's constants and buffers as locals, read once per tile: through the
        // descriptor the compiler re-reads them with scalar loads in the plane loop,
        // and every wait for one of those also waits for the LDS reads in flight
        const KParams kp = g.kp;
        const typename S::B bb = g.buf[par];
        if (q == g.nq - 1 && g.emit)
            march_tile<S, KW, MODE, true, false, TJ>(kp, bb, g.DI, tk, tj, seg, lds);
        else
            march_tile<S, KW, MODE, false, false, TJ>(kp, bb, g.DI, tk, tj, seg, lds);
    }
}

// Many grids, each with its own 256-thread tile shape (BatchGrid::tk: 64 for
// 64 x 4, 32 for 32 x 8), in blocks of 64 x 4 threads that run either shape's
// march by their linear thread id; staging sized for the larger (64 x 4).
template <class S, int MODE>
__global__ void __launch_bounds__(256, S::MINB)
k_march_batch_mix(const BatchGrid<typename S::B> *__restrict__ grids, const BatchTile *__restrict__ tiles,
                  uint32_t ntiles, uint32_t per_block, int q) {
    static_assert(march_lds_elems<S, 0, MODE, 8>() <= march_lds_elems<S, 1, MODE, 4>(), "staging of 32 x 8 fits");
    __shared__ typename S::T lds[march_lds_elems<S, 1, MODE, 4>()];
    const uint32_t b = logical_block(blockIdx.x, gridDim.x);
    const uint32_t t1 = min(ntiles, (b + 1) * per_block);
    for (uint32_t t = b * per_block; t < t1; t++) {
        const BatchTile bt = tiles[t];
        const BatchGrid<typename S::B> &g = grids[bt.grid];
        if (q >= g.nq) continue;
        const uint32_t tk = bt.tile % g.ntk, r = bt.tile / g.ntk;
        const uint32_t tj = r % g.ntj, seg = r / g.ntj;
        const int par = (g.cur0 + q) & 1;
        const bool emit = q == g.nq - 1 && g.emit;
        const KParams kp = g.kp;                  // locals, as in k_march_batch
        const typename S::B bb = g.buf[par];
        if (g.tk == 32) {
            if (emit) march_tile<S, 0, MODE, true, false, 8, 64>(kp, bb, g.DI, tk, tj, seg, lds);
            else march_tile<S, 0, MODE, false, false, 8, 64>(kp, bb, g.DI, tk, tj, seg, lds);
        } else {
            if (emit) march_tile<S, 1, MODE, true, false, 4, 64>(kp, bb, g.DI, tk, tj, seg, lds);
            else march_tile<S, 1, MODE, false, false, 4, 64>(kp, bb, g.DI, tk, tj, seg, lds);
        }
    }
}

template <class S, int MODE>
inline int launch_march_batch_mix_t(const BatchGrid<typename S::B> *grids, const BatchTile *tiles, uint32_t ntiles,
                                    uint32_t nblocks, uint32_t per_block, int q, hipStream_t s) {
    hipLaunchKernelGGL((k_march_batch_mix<S, MODE>), dim3(nblocks), dim3(64, 4), 0, s, grids, tiles, ntiles,
                       per_block, q);
    return (int)hipGetLastError();
}

template <class S, int MODE, int KW = 1, int TJ = MTJ>
inline int launch_march_batch_t(const BatchGrid<typename S::B> *grids, const BatchTile *tiles, uint32_t ntiles,
                                uint32_t nblocks, uint32_t per_block, int q, hipStream_t s) {
    hipLaunchKernelGGL((k_march_batch<S, KW, MODE, TJ>), dim3(nblocks), dim3(MG<KW, TJ>::TKM, TJ), 0, s, grids,
                       tiles, ntiles, per_block, q);
    return (int)hipGetLastError();
}

template <class S, int KW, int MODE>
inline int launch_march_t(bool emit, const KParams &kp, const typename S::B &b, int DI, hipStream_t s) {
    using G = MG<KW>;
    if (DI <= 0) DI = march_di(kp.N, KW, S::MARCH_BLOCKS);
    const uint32_t ntk = (kp.N + G::TKM - 1) / G::TKM, ntj = (kp.N + MTJ - 1) / MTJ, nseg = (kp.N + DI - 1) / DI;
    const dim3 grid(ntk * ntj * nseg), block(G::TKM, MTJ);
    if (emit) hipLaunchKernelGGL((k_march<S, KW, MODE, true>), grid, block, 0, s, kp, b, DI);
    else hipLaunchKernelGGL((k_march<S, KW, MODE, false>), grid, block, 0, s, kp, b, DI);
    return (int)hipGetLastError();
}

}  // namespace scp
