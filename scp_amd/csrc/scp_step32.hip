// scp_step32.hip — the FP32 compensated ("FP32C") form of the fused step.
//
// Storage: every scalar is a pair (hi, lo): hi is an FP32 word, lo the
// compensation word that carries what hi cannot represent (|lo| <= ulp(hi)/2),
// stored as FP32 (LO_F32, 8 bytes/scalar), bfloat16 (LO_BF16, 6 bytes), a
// signed byte in units of ulp(hi)/256 (LO_Q8, 5 bytes; LO_Q8P: on the
// positions only, 4.5 bytes on average) or not at all (LO_NONE, plain FP32). The stencil — forces, curls, covariant
// transport, plaquettes — is evaluated in FP32 on the hi words; the leapfrog
// accumulations (kicks, drift, E and link updates) are compensated FP32
// additions (leap32, scp_comp32.h), so increments far below ulp(x) accumulate
// in lo instead of being lost, and the state a run reaches does not depend on
// how it is split into calls. FP64 appears only in INIT (from the FP64 host
// state) and in the emitted snapshot state (hi + lo), which the FP64
// diagnostics (energy, Q, Gauss residual) read.
//
// Same schedule and z-march decomposition as the FP64 kernel (scp_step.hip).
#include "scp_comp32.h"
#include "scp_physics.h"

#include <hip/hip_runtime.h>

namespace scp {

// Compensation words of one site in one buffer (W, or P in / out) at hot index
// c, field f at f * fh + c: raw loads are decoded against their hi word (the
// LO_Q8 scale depends on it), stores encode against it. LO_Q8 words are packed
// four fields to a 32-bit word (LoQuads): load() reads the site's words before
// the field loop, store() writes them after it.
template <int LO, int NP> struct LoSite {
    using LW = LoWord<LO>;
    using W = typename LW::W;
    static constexpr bool PACKED = LO == LO_Q8;
    LoQuads in{}, out{};
    __device__ void load(const void *p, size_t fh, uint32_t c) {
        if constexpr (PACKED) {
#pragma unroll
            for (int q = 0; q < (NP + 3) / 4; q++)
                in.q[q] = __builtin_nontemporal_load((const uint32_t *)p + q * fh + c);
        }
    }
    __device__ float get(const void *p, int f, size_t fh, uint32_t c, float hi) const {
        if constexpr (PACKED) return LW::dec(in.get(f), hi);
        else if constexpr (LW::STORED) return LW::dec(__builtin_nontemporal_load((const W *)p + f * fh + c), hi);
        else return 0.0f;
    }
    __device__ void put(void *p, int f, size_t fh, uint32_t c, float hi, float lo) {
        if constexpr (PACKED) out.set(f, LW::enc(lo, hi));
        else if constexpr (LW::STORED) __builtin_nontemporal_store(LW::enc(lo, hi), (W *)p + f * fh + c);
    }
    __device__ void store(void *p, size_t fh, uint32_t c) const {
        if constexpr (PACKED) {
#pragma unroll
            for (int q = 0; q < (NP + 3) / 4; q++)
                __builtin_nontemporal_store(out.q[q], (uint32_t *)p + q * fh + c);
        }
    }
};

template <int MODE, int KIND, bool EMIT, int LO>
__global__ void __launch_bounds__(TK * TJ)
k_step32(KParams kp, Bufs32 b, int DI) {
    constexpr int NP = MODE == 2 ? 15 : (MODE == 1 ? 12 : 6);
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
        else {
            const uint32_t ipx = IX(ip1, j, k), imx = IX(im1, j, k);
            const auto nbi = [&](int d, int sg) -> uint32_t {
                if (d == 3) return ci;
                if (d == 0) return sg > 0 ? ipx : imx;
                if (d == 1) return sg > 0 ? jp : jm;
                return sg > 0 ? kpx : kmx;
            };
            forces_real(kp, c0, cp, cm, nb, AUX, nbi, acc);
        }

        if constexpr (KIND == K_FORCES) {
#pragma unroll
            for (int f = 0; f < NP; f++) b.acc[(size_t)f * fs + cc] = (double)acc[f];
        } else if constexpr (KIND == K_INIT) {
            // first half kick and drift in FP64 from the FP64 state, then split
            LoSite<lo_kind_w(LO), NP> lw;
            LoSite<lo_kind_p(LO), NP> lp;
#pragma unroll
            for (int f = 0; f < NP; f++) {
                const size_t o = (size_t)f * fh + chh, oc = (size_t)f * fs + cc;
                const double a = kp.init_acc ? b.acc[oc] : (double)acc[f];   // a given a_n, or F(x)
                double w = b.vo[oc] + kp.hdt * a;
                double x = b.xo[oc], xn;
                if (MODE == 2 && f >= F_TH) {
                    double t = x + kp.gad * w;
                    xn = t - 2.0 * SCP_PI * rint(t / (2.0 * SCP_PI));
                } else {
                    xn = x + kp.dt * w;
                }
                float wh = (float)w, xh = (float)xn;
                b.w[o] = wh;
                lw.put(b.wlo, f, fh, chh, wh, (float)(w - (double)wh));
                b.pout[o] = xh;
                lp.put(b.poutlo, f, fh, chh, xh, (float)(xn - (double)xh));
            }
            lw.store(b.wlo, fh, chh);
            lp.store(b.poutlo, fh, chh);
        } else {
            float dmp = 1.0f;
            bool damp = false;
            if (kp.damp) {
                double dd;
                damp = sponge(kp, i, j, k, dd);
                dmp = (float)dd;
            }
            LoSite<lo_kind_w(LO), NP> lw;
            LoSite<lo_kind_p(LO), NP> lp;
            lw.load(b.wlo, fh, chh);
            lp.load(b.pinlo, fh, chh);
#pragma unroll
            for (int f = 0; f < NP; f++) {
                const size_t o = (size_t)f * fh + chh, oc = (size_t)f * fs + cc;
                float wh = __builtin_nontemporal_load(&b.w[o]), wl = lw.get(b.wlo, f, fh, chh, wh);
                float xh = c0[f], xl = lp.get(b.pinlo, f, fh, chh, xh), vh, vl;
                const float xh0 = xh, xl0 = xl;
                leap32(MODE == 2 && f >= F_TH, acc[f], damp, dmp, hdt, dt, gad, wh, wl, xh, xl, vh, vl);
                __builtin_nontemporal_store(wh, &b.w[o]);
                lw.put(b.wlo, f, fh, chh, wh, wl);
                __builtin_nontemporal_store(xh, &b.pout[o]);
                lp.put(b.poutlo, f, fh, chh, xh, xl);
                if constexpr (EMIT) {
                    __builtin_nontemporal_store((double)xh0 + (double)xl0, &b.xo[oc]);
                    __builtin_nontemporal_store((double)vh + (double)vl, &b.vo[oc]);
                }
            }
            lw.store(b.wlo, fh, chh);
            lp.store(b.poutlo, fh, chh);
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
    if (lo == LO_Q8) return launch32<MODE, LO_Q8>(kind, emit, kp, b, s);
    if (lo == LO_Q8P) return launch32<MODE, LO_Q8P>(kind, emit, kp, b, s);
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
