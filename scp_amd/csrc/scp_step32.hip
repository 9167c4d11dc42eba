// scp_step32.hip — the FP32 compensated ("FP32C") form of the fused step.
//
// Storage: every scalar is a pair (hi, lo): hi is an FP32 word, lo the
// compensation word that carries what hi cannot represent (|lo| <= ulp(hi)/2),
// stored as FP32 (LO_F32, 8 bytes/scalar), bfloat16 (LO_BF16, 6 bytes) or not
// at all (LO_NONE, plain FP32 for comparison). The stencil — forces, curls,
// covariant transport, plaquettes — is evaluated in FP32 on the hi words; the
// leapfrog accumulations v += h a, x += dt v, E and link updates are
// compensated FP32 additions (TwoSum with the increment formed by FMAs), so
// increments far below ulp(x) accumulate in lo instead of being lost. Both half
// kicks are taken every step, so results do not depend on how a run is split
// into calls. FP64 appears
// only in INIT (from the FP64 host state) and in the emitted snapshot state
// (hi + lo), which the FP64 diagnostics (energy, Q, Gauss residual) read.
//
// Same schedule and z-march decomposition as the FP64 kernel (scp_step.hip).
// Compiled with FMA contraction allowed: the compensated sums contain no
// products (the increment is formed by an explicit fmaf), so contraction
// cannot change them.
#include "scp_physics.h"

#include <hip/hip_runtime.h>

namespace scp {

constexpr float TWO_PI_HI = 6.28318548202514648f;      // fl32(2 pi)
constexpr float TWO_PI_LO = -1.7484555314695172e-07f;  // 2 pi - TWO_PI_HI
constexpr float INV_TWO_PI = 0.159154943091895336f;

__device__ inline void fast_two_sum(float a, float b, float &s, float &e) {
    s = a + b;
    e = b - (s - a);
}
// A constant as an unevaluated pair c = h + l of FP32 words (the FP64 step
// constants dt, dt/2, -g dx dt split once), so increments are formed with the
// FP64 constant and not with its FP32 rounding.
struct DF {
    float h, l;
};
__device__ inline DF df_split(double c) {
    const float h = (float)c;
    return {h, (float)(c - (double)h)};
}

// (hi, lo) += c * (xh + xl): the increment is formed with FMAs and added by
// Knuth's TwoSum, which is exact for any magnitudes (velocities and fields
// pass through zero, so |hi| >= |increment| cannot be assumed).
__device__ inline void comp_add(float &hi, float &lo, DF c, float xh, float xl) {
    const float y = fmaf(c.h, xh, fmaf(c.l, xh, fmaf(c.h, xl, lo)));
    const float t = hi + y;
    const float bp = t - hi;
    lo = (hi - (t - bp)) + (y - bp);
    hi = t;
}
// (hi, lo) *= d
__device__ inline void df_scale(float &hi, float &lo, float d) {
    float p = __fmul_rn(hi, d);
    float e = fmaf(hi, d, -p) + __fmul_rn(lo, d);
    fast_two_sum(p, e, hi, lo);
}

template <int LO> struct LoStore;
template <> struct LoStore<LO_F32> {
    __device__ static float get(const void *p, size_t o) { return __builtin_nontemporal_load((const float *)p + o); }
    __device__ static void put(void *p, size_t o, float v) { __builtin_nontemporal_store(v, (float *)p + o); }
};
template <> struct LoStore<LO_BF16> {
    __device__ static float get(const void *p, size_t o) {
        uint32_t b = (uint32_t)__builtin_nontemporal_load((const uint16_t *)p + o) << 16;
        return __uint_as_float(b);
    }
    __device__ static void put(void *p, size_t o, float v) {      // round to nearest even
        uint32_t u = __float_as_uint(v);
        u += 0x7FFFu + ((u >> 16) & 1u);
        __builtin_nontemporal_store((uint16_t)(u >> 16), (uint16_t *)p + o);
    }
};
template <> struct LoStore<LO_NONE> {
    __device__ static float get(const void *, size_t) { return 0.0f; }
    __device__ static void put(void *, size_t, float) {}
};

template <int MODE, int KIND, bool EMIT, int LO>
__global__ void __launch_bounds__(TK * TJ)
k_step32(KParams kp, Bufs32 b, int DI) {
    constexpr int NP = MODE == 2 ? 15 : (MODE == 1 ? 12 : 6);
    using LS = LoStore<LO>;
    const int N = kp.N;
    const Geo g = geometry(N, DI);
    if (!g.valid || g.i0 >= N) return;
    // stencil source: FP32 hi words of P (padded rows, kp.pitch), or the FP64
    // emitted state rounded (INIT / FORCES, unpadded); outputs P', W and their
    // words hot, xo/vo/acc cold
    const uint32_t fs = kp.fs, fh = kp.fsh;
    constexpr bool hot = (KIND == K_STEP);
    const uint32_t RS = hot ? kp.pitch : (uint32_t)N, PS = hot ? kp.plane : (uint32_t)N * N;
    auto X = [&](int f, uint32_t idx) -> float {
        if constexpr (KIND == K_STEP) return b.pin[(size_t)f * fh + idx];
        else return (float)b.xo[(size_t)f * fs + idx];
    };
    auto IX = [&](int a, int bb, int c) -> uint32_t { return (uint32_t)a * PS + (uint32_t)bb * RS + (uint32_t)c; };
    auto AUX = [&](int f, uint32_t idx) -> float { return (float)b.aux[(size_t)f * fs + idx]; };   // real mode: hot == cold
    const int j = g.j, k = g.k;
    const int jp1 = wp(j, N), jm1 = wm(j, N), kp1 = wp(k, N), km1 = wm(k, N);

    float cm[NP], c0[NP], cp[NP];
    {
        const uint32_t a = IX(wm(g.i0, N), j, k), c = IX(g.i0, j, k);
#pragma unroll
        for (int f = 0; f < NP; f++) { cm[f] = X(f, a); c0[f] = X(f, c); }
    }
    GaugeCarry<float> cy{};
    if constexpr (MODE == 2) {
        const int im0 = wm(g.i0, N);
        gauge_prologue(cm[F_TH + 0], cm[F_TH + 1], cm[F_TH + 2], c0[F_TH + 1], c0[F_TH + 2],
                       X(F_TH + 0, IX(im0, jp1, k)), X(F_TH + 0, IX(im0, j, kp1)), cy);
    }
    const DF hdt = df_split(kp.hdt), dt = df_split(kp.dt), gad = df_split(kp.gad);

    for (int i = g.i0; i < g.i1; i++) {
        const int ip1 = wp(i, N), im1 = wm(i, N);
        const uint32_t ci = IX(i, j, k);                          // source index
        const uint32_t cc = sidx(N, i, j, k);                     // cold (xo, vo, acc)
        const uint32_t chh = (uint32_t)i * kp.plane + (uint32_t)j * kp.pitch + (uint32_t)k;   // hot (P', W, words)
        {
            const uint32_t a = IX(ip1, j, k);
#pragma unroll
            for (int f = 0; f < NP; f++) cp[f] = X(f, a);
        }
        const uint32_t jp = IX(i, jp1, k), jm = IX(i, jm1, k);
        const uint32_t kpx = IX(i, j, kp1), kmx = IX(i, j, km1);
        float acc[NP];
        const GlobalNB<float, decltype(X)> nb{X, N, i, j, k, jp, jm, kpx, kmx, RS, PS};
        if constexpr (MODE == 2)
            forces_gauge(kp, c0, cp, cm, nb, cy, acc);
        else if constexpr (MODE == 1)
            forces_complex(kp, c0, cp, cm, nb, acc);
        else
            forces_real(kp, c0, cp, cm, X, AUX, ci, IX(ip1, j, k), IX(im1, j, k), jp, jm,
                        kpx, kmx, acc);

        if constexpr (KIND == K_FORCES) {
#pragma unroll
            for (int f = 0; f < NP; f++) b.acc[(size_t)f * fs + cc] = (double)acc[f];
        } else if constexpr (KIND == K_INIT) {
            // first half kick and drift in FP64 from the FP64 state, then split
#pragma unroll
            for (int f = 0; f < NP; f++) {
                const size_t o = (size_t)f * fh + chh, oc = (size_t)f * fs + cc;
                double w = b.vo[oc] + kp.hdt * (double)acc[f];
                double x = b.xo[oc], xn;
                if (MODE == 2 && f >= F_TH) {
                    double t = x + kp.gad * w;
                    xn = t - 2.0 * SCP_PI * rint(t / (2.0 * SCP_PI));
                } else {
                    xn = x + kp.dt * w;
                }
                float wh = (float)w, xh = (float)xn;
                b.w[o] = wh;
                LS::put(b.wlo, o, (float)(w - (double)wh));
                b.pout[o] = xh;
                LS::put(b.poutlo, o, (float)(xn - (double)xh));
            }
        } else {
            float dmp = 1.0f;
            bool damp = false;
            if (kp.damp) {
                double dd;
                damp = sponge(kp, i, j, k, dd);
                dmp = (float)dd;
            }
#pragma unroll
            for (int f = 0; f < NP; f++) {
                const size_t o = (size_t)f * fh + chh, oc = (size_t)f * fs + cc;
                float wh = __builtin_nontemporal_load(&b.w[o]), wl = LS::get(b.wlo, o);
                float vh = 0.0f, vl = 0.0f;
                // both half kicks, every step (also between emitted steps): the result then
                // does not depend on how a run is split into calls or batch chains
                vh = wh; vl = wl;
                comp_add(vh, vl, hdt, acc[f], 0.0f);                       // stage 4: v = W + dt/2 a
                if (damp) df_scale(vh, vl, dmp);                 // stage 5
                wh = vh; wl = vl;
                comp_add(wh, wl, hdt, acc[f], 0.0f);                       // next stage 1
                float xh = c0[f], xl = LS::get(b.pinlo, o);
                const float xh0 = xh, xl0 = xl;
                if (MODE == 2 && f >= F_TH) {                       // link drift + wrap
                    comp_add(xh, xl, gad, wh, wl);
                    float n = rintf(xh * INV_TWO_PI);
                    if (n != 0.0f) {
                        comp_add(xh, xl, DF{-n, 0.0f}, TWO_PI_HI, TWO_PI_LO);
                    }
                } else {
                    comp_add(xh, xl, dt, wh, wl);                     // next stage 2
                }
                __builtin_nontemporal_store(wh, &b.w[o]);
                LS::put(b.wlo, o, wl);
                __builtin_nontemporal_store(xh, &b.pout[o]);
                LS::put(b.poutlo, o, xl);
                if constexpr (EMIT) {
                    __builtin_nontemporal_store((double)xh0 + (double)xl0, &b.xo[oc]);
                    __builtin_nontemporal_store((double)vh + (double)vl, &b.vo[oc]);
                }
            }
        }
#pragma unroll
        for (int f = 0; f < NP; f++) { cm[f] = c0[f]; c0[f] = cp[f]; }
    }
}

// FP32C form of k_real_aux (scp_step.hip)
__global__ void k_real_aux32(KParams kp, const float *pin, const float *vel, double *aux) {
    const int N = kp.N;
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= kp.n3) return;
    const uint32_t fs = kp.fs;
    int k = c % N, j = (c / N) % N, i = c / ((uint32_t)N * N);
    uint32_t np[3] = {sidx(N, wp(i, N), j, k), sidx(N, i, wp(j, N), k), sidx(N, i, j, wp(k, N))};
    uint32_t nm[3] = {sidx(N, wm(i, N), j, k), sidx(N, i, wm(j, N), k), sidx(N, i, j, wm(k, N))};
    auto F = [&](int f, uint32_t idx) -> float { return pin[(size_t)f * fs + idx]; };
    const float h1 = (float)kp.h1;
    float cp[3];
    cp[0] = (F(2, np[1]) - F(2, nm[1]) - F(1, np[2]) + F(1, nm[2])) * h1;
    cp[1] = (F(0, np[2]) - F(0, nm[2]) - F(2, np[0]) + F(2, nm[0])) * h1;
    cp[2] = (F(1, np[0]) - F(1, nm[0]) - F(0, np[1]) + F(0, nm[1])) * h1;
    if (kp.ACS != 0)
        for (int a = 0; a < 3; a++) aux[(size_t)a * fs + c] = cp[a] * 0.5f - F(3 + a, c);
    if (kp.BH != 0) {
        float t0 = F(3, c), t1 = F(4, c), t2 = F(5, c);
        float coeff = 0.5f * (float)kp.BH * (t0 * t0 + t1 * t1 + t2 * t2);
        for (int a = 0; a < 3; a++) aux[(size_t)(3 + a) * fs + c] = coeff * cp[a];
    }
    if (kp.SFREQ > 0) {
        auto V = [&](int f) -> float { return vel[(size_t)f * fs + c]; };
        const float tv = V(3) * V(3) + V(4) * V(4) + V(5) * V(5);
        const float th = F(3, c) * F(3, c) + F(4, c) * F(4, c) + F(5, c) * F(5, c);
        aux[(size_t)6 * fs + c] = (double)(sqrtf(tv) / (sqrtf(th) + 0.01f));
    }
}

static inline int pick_di32(int N) {
    uint32_t tiles = ((N + TK - 1) / TK) * ((N + TJ - 1) / TJ);
    int di = N;
    while (di > 4 && (uint64_t)tiles * ((N + di - 1) / di) < 2048) di /= 2;
    return di;
}

template <int MODE, int LO>
static int launch32(int kind, bool emit, const KParams &kp, const Bufs32 &b, hipStream_t s) {
    const int DI = pick_di32(kp.N);
    uint32_t ntk = (kp.N + TK - 1) / TK, ntj = (kp.N + TJ - 1) / TJ, nseg = (kp.N + DI - 1) / DI;
    dim3 grid(ntk * ntj * nseg), block(TK, TJ);
    if (kind == K_INIT) hipLaunchKernelGGL((k_step32<MODE, K_INIT, false, LO>), grid, block, 0, s, kp, b, DI);
    else if (kind == K_FORCES) hipLaunchKernelGGL((k_step32<MODE, K_FORCES, false, LO>), grid, block, 0, s, kp, b, DI);
    else if (emit) hipLaunchKernelGGL((k_step32<MODE, K_STEP, true, LO>), grid, block, 0, s, kp, b, DI);
    else hipLaunchKernelGGL((k_step32<MODE, K_STEP, false, LO>), grid, block, 0, s, kp, b, DI);
    return (int)hipGetLastError();
}

template <int MODE>
static int launch32m(int kind, bool emit, int lo, const KParams &kp, const Bufs32 &b, hipStream_t s) {
    if (lo == LO_F32) return launch32<MODE, LO_F32>(kind, emit, kp, b, s);
    if (lo == LO_BF16) return launch32<MODE, LO_BF16>(kind, emit, kp, b, s);
    return launch32<MODE, LO_NONE>(kind, emit, kp, b, s);
}

int launch_step_f32c(int mode, int kind, bool emit, int lo_kind, const KParams &kp, const Bufs32 &b,
                     hipStream_t s) {
    if (mode == 2) return launch32m<2>(kind, emit, lo_kind, kp, b, s);
    if (mode == 1) return launch32m<1>(kind, emit, lo_kind, kp, b, s);
    return launch32m<0>(kind, emit, lo_kind, kp, b, s);
}

int launch_real_aux_f32(const KParams &kp, const float *pin, const float *vel, double *aux, hipStream_t s) {
    hipLaunchKernelGGL(k_real_aux32, dim3((kp.n3 + 255) / 256), dim3(256), 0, s, kp, pin, vel, aux);
    return (int)hipGetLastError();
}

}  // namespace scp
