// scp_step.hip — gfx950 kernels of the sfa/sim leapfrog step.
//
// One fused launch per step. The stored state between launches is
//   P = x_{n+1}              positions already drifted for the next step
//   W = v_n + (dt/2) a_n     the half-kicked velocities (E for the links)
// so a launch reads P (with neighbours) and W (own site only), evaluates
// a_{n+1} = F(P), completes step n+1 (v_{n+1} = W + dt/2 a, sponge), starts
// step n+2 (W' = v_{n+1} + dt/2 a, P' = P + dt W', links wrapped), and writes
// P' and W'. Every floating-point operation is the reference's own, in its
// order (scp_sim.c:1583-1716), so with -ffp-contract=off the FP64 path rounds
// exactly like the reference CPU path; no acceleration arrays are stored.
//
// Work decomposition: a block is a 64(k) x 4(j) tile of one-site-per-thread
// columns that marches along i over a segment of DI planes, keeping the
// planes i-1, i, i+1 of its own column in registers. j/k neighbours are read
// from global memory; the blocks of adjacent j tiles run together on the same
// XCD (XCD-aware block numbering), so those reads hit that XCD's L2.
#include "scp_physics.h"

#include <hip/hip_runtime.h>

namespace scp {

// ---------------------------------------------------------------------------
// the fused step kernel, FP64 storage
// ---------------------------------------------------------------------------

// Gradient-pinned boundary (bc 1, apply_gradient_bc, scp_sim.c:1480-1555), the
// part a site can do alone: what the y-edge pass needs from this launch — the
// edge rows' accelerations (the edges keep them) and the velocities of the two
// rows inside each edge (the x-slab sites are pinned in the kernel itself).
__device__ inline void bc1_side(const KParams &kp, const Bufs<double> &b, int f, int i, int j, int k, double a,
                                double v) {
    const int N = kp.N;
    const size_t NN = (size_t)N * N, base = (size_t)f * 6 * NN + (size_t)i * N + k;
    if (j == 0) b.side[base] = a;
    if (j == N - 1) b.side[base + NN] = a;
    if (j == 1) b.side[base + 2 * NN] = v;
    if (j == 2) b.side[base + 3 * NN] = v;
    if (j == N - 2) b.side[base + 4 * NN] = v;
    if (j == N - 3) b.side[base + 5 * NN] = v;
}

template <int MODE, int KIND, bool EMIT>
__global__ void __launch_bounds__(TK * TJ)
k_step64(KParams kp, Bufs<double> b, int DI) {
    constexpr int NP = MODE == 2 ? 15 : (MODE == 1 ? 12 : 6);
    const int N = kp.N;
    const Geo g = geometry(N, DI);
    if (!g.valid || g.i0 >= N) return;
    // the stencil source: P (padded rows, kp.pitch) when stepping, the emitted
    // state xo (unpadded) at INIT / FORCES; outputs P', W hot, xo/vo/acc cold
    const uint32_t fs = kp.fs, fh = kp.fsh;
    const bool hot = (KIND == K_STEP);
    const uint32_t RS = hot ? kp.pitch : (uint32_t)N, PS = hot ? kp.plane : (uint32_t)N * N, SF = hot ? fh : fs;
    const double *__restrict__ src = hot ? b.pin : b.xo;
    auto X = [&](int f, uint32_t idx) -> double { return src[(size_t)f * SF + idx]; };
    auto IX = [&](int a, int bb, int c) -> uint32_t { return (uint32_t)a * PS + (uint32_t)bb * RS + (uint32_t)c; };
    auto AUX = [&](int f, uint32_t idx) -> double { return b.aux[(size_t)f * fs + idx]; };   // real mode: hot == cold
    const int j = g.j, k = g.k;
    const int jp1 = wp(j, N), jm1 = wm(j, N), kp1 = wp(k, N), km1 = wm(k, N);

    double cm[NP], c0[NP], cp[NP];
    {
        const int im1 = wm(g.i0, N);
        const uint32_t a = IX(im1, j, k), c = IX(g.i0, j, k);
#pragma unroll
        for (int f = 0; f < NP; f++) { cm[f] = X(f, a); c0[f] = X(f, c); }
    }
    GaugeCarry<double> cy{};
    if constexpr (MODE == 2) {
        const int im0 = wm(g.i0, N);
        gauge_prologue(cm[F_TH + 0], cm[F_TH + 1], cm[F_TH + 2], c0[F_TH + 1], c0[F_TH + 2],
                       X(F_TH + 0, IX(im0, jp1, k)), X(F_TH + 0, IX(im0, j, kp1)), cy);
    }

    for (int i = g.i0; i < g.i1; i++) {
        const int ip1 = wp(i, N), im1 = wm(i, N);
        const uint32_t ci = IX(i, j, k);                          // source index
        const uint32_t cc = sidx(N, i, j, k);                     // cold (xo, vo, acc)
        const uint32_t chh = (uint32_t)i * kp.plane + (uint32_t)j * kp.pitch + (uint32_t)k;   // hot (P', W)
        {
            const uint32_t a = IX(ip1, j, k);
#pragma unroll
            for (int f = 0; f < NP; f++) cp[f] = X(f, a);
        }
        const uint32_t jp = IX(i, jp1, k), jm = IX(i, jm1, k);
        const uint32_t kpx = IX(i, j, kp1), kmx = IX(i, j, km1);
        double acc[NP];
        const GlobalNB<double, decltype(X)> nb{X, N, i, j, k, jp, jm, kpx, kmx, RS, PS};
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
            for (int f = 0; f < NP; f++) b.acc[(size_t)f * fs + cc] = acc[f];
        } else {
            double dmp = 1.0;
            const bool damp = (KIND == K_STEP) && kp.damp && sponge(kp, i, j, k, dmp);
            // bc 1: the x-slabs (i < M or i >= N - M) return to the pinned state with
            // zero acceleration (stage 5, C:1484-1508)
            const bool pinned = (KIND == K_STEP && MODE != 2) && kp.bc1 && (i < kp.gm || i >= N - kp.gm);
#pragma unroll
            for (int f = 0; f < NP; f++) {
                const size_t o = (size_t)f * fh + chh, oc = (size_t)f * fs + cc;
                double v, a = acc[f];
                double x = c0[f];
                if constexpr (KIND == K_INIT) {
                    v = b.vo[oc];
                    if (kp.init_acc) a = b.acc[oc];        // a_n kept across a rescale of x_n, v_n
                } else {
                    v = __builtin_nontemporal_load(&b.w[o]);
                    v += kp.hdt * a;                       // stage 4 (C:1680-1696)
                    if (damp) v *= dmp;                    // stage 5 (C:1454-1460)
                }
                if constexpr (KIND == K_STEP && MODE != 2) {
                    if (kp.bc1) {                          // stage 5, bc 1
                        if (pinned) {
                            x = b.pinx[oc];
                            v = b.pinv[oc];
                            a = 0.0;
                        } else {
                            bc1_side(kp, b, f, i, j, k, a, v);
                        }
                    }
                }
                double w = v + kp.hdt * a;                 // next stage 1 (C:1597-1613)
                double xn;
                if (MODE == 2 && f >= F_TH) {              // next stage 2, links (C:1641-1651)
                    double t = x + kp.gad * w;
                    xn = t - 2.0 * SCP_PI * rint(t / (2.0 * SCP_PI));
                } else {
                    xn = x + kp.dt * w;                    // next stage 2 (C:1629-1638)
                }
                __builtin_nontemporal_store(w, &b.w[o]);
                __builtin_nontemporal_store(xn, &b.pout[o]);
                if constexpr (EMIT && KIND == K_STEP) {
                    __builtin_nontemporal_store(x, &b.xo[oc]);
                    __builtin_nontemporal_store(v, &b.vo[oc]);
                }
            }
        }
#pragma unroll
        for (int f = 0; f < NP; f++) { cm[f] = c0[f]; c0[f] = cp[f]; }
    }
}

// Real-mode pass 1 (C:755-786): Cosserat mismatch and hardening vector; and
// the local theta frequency |theta_vel| / (|theta| + 0.01) of the
// frequency-mismatch term (scp_sim.cu:1204-1217), from the velocities the
// force sees (the half-kicked ones in a step). A pass of its own because a
// fused step updates the velocities in place while neighbours read them.
template <typename T>
__global__ void k_real_aux(KParams kp, const T *pin, const T *vel, double *aux) {
    const int N = kp.N;
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= kp.n3) return;
    const uint32_t fs = kp.fs;
    int k = c % N, j = (c / N) % N, i = c / ((uint32_t)N * N);
    uint32_t np[3] = {sidx(N, wp(i, N), j, k), sidx(N, i, wp(j, N), k), sidx(N, i, j, wp(k, N))};
    uint32_t nm[3] = {sidx(N, wm(i, N), j, k), sidx(N, i, wm(j, N), k), sidx(N, i, j, wm(k, N))};
    auto F = [&](int f, uint32_t idx) -> double { return (double)pin[(size_t)f * fs + idx]; };
    const double h1 = kp.h1;
    double cp[3];
    cp[0] = (F(2, np[1]) - F(2, nm[1]) - F(1, np[2]) + F(1, nm[2])) * h1;
    cp[1] = (F(0, np[2]) - F(0, nm[2]) - F(2, np[0]) + F(2, nm[0])) * h1;
    cp[2] = (F(1, np[0]) - F(1, nm[0]) - F(0, np[1]) + F(0, nm[1])) * h1;
    if (kp.ACS != 0)
        for (int a = 0; a < 3; a++) aux[(size_t)a * fs + c] = cp[a] * 0.5 - F(3 + a, c);
    if (kp.BH != 0) {
        double t0 = F(3, c), t1 = F(4, c), t2 = F(5, c);
        double T2 = t0 * t0 + t1 * t1 + t2 * t2;
        double coeff = 0.5 * kp.BH * T2;
        for (int a = 0; a < 3; a++) aux[(size_t)(3 + a) * fs + c] = coeff * cp[a];
    }
    if (kp.SFREQ > 0) {
        auto V = [&](int f) -> double { return (double)vel[(size_t)f * fs + c]; };
        const double tv = V(3) * V(3) + V(4) * V(4) + V(5) * V(5);
        const double th = F(3, c) * F(3, c) + F(4, c) * F(4, c) + F(5, c) * F(5, c);
        aux[(size_t)6 * fs + c] = sqrt(tv) / (sqrt(th) + 0.01);
    }
}

static inline dim3 step_grid(int N, int DI) {
    uint32_t ntk = (N + TK - 1) / TK, ntj = (N + TJ - 1) / TJ, nseg = (N + DI - 1) / DI;
    return dim3(ntk * ntj * nseg);
}

// segment depth: long enough to amortise the 2-plane prologue, short enough
// to give >= ~4 blocks per CU
static inline int pick_di(int N) {
    uint32_t tiles = ((N + TK - 1) / TK) * ((N + TJ - 1) / TJ);
    int di = N;
    while (di > 4 && (uint64_t)tiles * ((N + di - 1) / di) < 2048) di /= 2;
    return di;
}

template <int MODE>
static int launch64(int kind, bool emit, const KParams &kp, const Bufs<double> &b, hipStream_t s) {
    const int DI = pick_di(kp.N);
    dim3 grid = step_grid(kp.N, DI), block(TK, TJ);
    if (kind == K_INIT) hipLaunchKernelGGL((k_step64<MODE, K_INIT, false>), grid, block, 0, s, kp, b, DI);
    else if (kind == K_FORCES) hipLaunchKernelGGL((k_step64<MODE, K_FORCES, false>), grid, block, 0, s, kp, b, DI);
    else if (emit) hipLaunchKernelGGL((k_step64<MODE, K_STEP, true>), grid, block, 0, s, kp, b, DI);
    else hipLaunchKernelGGL((k_step64<MODE, K_STEP, false>), grid, block, 0, s, kp, b, DI);
    return (int)hipGetLastError();
}

int launch_step_f64(int mode, int kind, bool emit, const KParams &kp, const Bufs<double> &b,
                    hipStream_t s) {
    if (mode == 2) return launch64<2>(kind, emit, kp, b, s);
    if (mode == 1) return launch64<1>(kind, emit, kp, b, s);
    return launch64<0>(kind, emit, kp, b, s);
}

// Gradient-pinned boundary, y edges (apply_gradient_bc, C:1510-1553): for i in
// [M, N-M) the rows j = 0 and j = N-1 of every field are extrapolated linearly
// from the two rows inside, positions and velocities alike (x_0 = 2 x_1 - x_2);
// an edge keeps the acceleration of its own force evaluation. Runs after the
// step launch, whose stencil source (b.pin) still holds the inside rows'
// positions and whose side rows hold their velocities and the edges'
// accelerations; recomputes the edges' W and P' (next stages 1-2) and the
// emitted state. One thread per (field, edge, i, k).
template <bool EMIT>
__global__ void k_bc1_edges(KParams kp, Bufs<double> b, int nf) {
    const int N = kp.N, M = kp.gm, NI = N - 2 * M;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (NI <= 0 || t >= (uint32_t)nf * 2u * (uint32_t)NI * (uint32_t)N) return;
    const int k = t % N;
    uint32_t r = t / N;
    const int i = M + (int)(r % NI);
    r /= NI;
    const int e = r % 2, f = r / 2;
    const int j = e ? N - 1 : 0, j1 = e ? N - 2 : 1, j2 = e ? N - 3 : 2;
    const size_t NN = (size_t)N * N, base = (size_t)f * 6 * NN + (size_t)i * N + k;
    const size_t fb = (size_t)f * kp.fsh + (size_t)i * kp.plane + k;
    const double x1 = b.pin[fb + (size_t)j1 * kp.pitch], x2 = b.pin[fb + (size_t)j2 * kp.pitch];
    const double v1 = b.side[base + (e ? 4 : 2) * NN], v2 = b.side[base + (e ? 5 : 3) * NN];
    const double a = b.side[base + e * NN];
    const double x = 2 * x1 - x2;
    const double v = 2 * v1 - v2;
    const double w = v + kp.hdt * a;                       // next stage 1
    const double xn = x + kp.dt * w;                       // next stage 2
    const size_t o = fb + (size_t)j * kp.pitch;
    b.w[o] = w;
    b.pout[o] = xn;
    if constexpr (EMIT) {
        const size_t oc = (size_t)f * kp.fs + (size_t)i * NN + (size_t)j * N + k;
        b.xo[oc] = x;
        b.vo[oc] = v;
    }
}

int launch_bc1_edges(bool emit, int nf, const KParams &kp, const Bufs<double> &b, hipStream_t s) {
    const int NI = kp.N - 2 * kp.gm;
    if (NI <= 0) return 0;
    const uint32_t n = (uint32_t)nf * 2u * (uint32_t)NI * (uint32_t)kp.N;
    if (emit) hipLaunchKernelGGL((k_bc1_edges<true>), dim3((n + 255) / 256), dim3(256), 0, s, kp, b, nf);
    else hipLaunchKernelGGL((k_bc1_edges<false>), dim3((n + 255) / 256), dim3(256), 0, s, kp, b, nf);
    return (int)hipGetLastError();
}

int launch_real_aux(const KParams &kp, const double *pin, const double *vel, double *aux, hipStream_t s) {
    hipLaunchKernelGGL((k_real_aux<double>), dim3((kp.n3 + 255) / 256), dim3(256), 0, s, kp, pin, vel, aux);
    return (int)hipGetLastError();
}

}  // namespace scp
