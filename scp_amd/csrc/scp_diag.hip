// scp_diag.hip — FP64 diagnostics and Gauss-law kernels.
//
// Diagnostics (reference scp_sim.c:1722-2243) are one fused pass over the
// emitted state: every energy term, Noether charges, P_int, theta_rms, the
// Gauss residual statistics, E_em and Q_flux accumulate per thread, reduce per
// block with a fixed butterfly, and a single-block kernel sums the block
// partials in block order. The result is bitwise reproducible run to run
// (the reference's OpenMP/atomicAdd sums are not; they agree to a few ulps).
#include "scp_diag_site.h"
#include "scp_kernels.h"

#include <hip/hip_runtime.h>

namespace scp {

__device__ inline uint32_t sidx3(int N, int i, int j, int k) {
    return ((uint32_t)i * (uint32_t)N + (uint32_t)j) * (uint32_t)N + (uint32_t)k;
}
__device__ inline int wpd(int a, int N) { return a + 1 == N ? 0 : a + 1; }
__device__ inline int wmd(int a, int N) { return a == 0 ? N - 1 : a - 1; }

// two blocks per CU for pass 1: the compiler caps it at 256 registers (a few
// spills) and the second block hides the neighbour-load latency (measured at
// 512^3 gauged: 23.8 ms against 29.9 ms with one block and 49.5 ms with three)
#ifndef DIAG_MINB
#define DIAG_MINB 2
#endif

// deterministic block reduction of NV values per thread into out[0..NV)
template <int NV, bool MAXV>
__device__ inline void block_reduce(double (&v)[NV], double *out) {
    __shared__ double red[DIAG_BLOCK / 64][NV];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NV; q++) {
        double x = v[q];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            double y = __shfl_xor(x, o, 64);
            x = MAXV ? fmax(x, y) : x + y;
        }
        v[q] = x;
    }
    if (lane == 0)
#pragma unroll
        for (int q = 0; q < NV; q++) red[wv][q] = v[q];
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NV; q++) {
            double x = red[0][q];
            for (int w = 1; w < DIAG_BLOCK / 64; w++) x = MAXV ? fmax(x, red[w][q]) : x + red[w][q];
            out[q] = x;
        }
    }
    __syncthreads();
}

template <int MODE>
__global__ void __launch_bounds__(DIAG_BLOCK, DIAG_MINB)
k_diag1(KParams kp, DiagParams dp, const double *__restrict__ xo, const double *__restrict__ vo,
        double *partials) {
    const int N = kp.N;
    const uint32_t fs = kp.fs, n3 = kp.n3;
    auto X = [&](int f, uint32_t c) -> double { return xo[(size_t)f * fs + c]; };
    auto V = [&](int f, uint32_t c) -> double { return vo[(size_t)f * fs + c]; };
    double S[DS_NSUM], M[DM_NMAX];
#pragma unroll
    for (int q = 0; q < DS_NSUM; q++) S[q] = 0.0;
#pragma unroll
    for (int q = 0; q < DM_NMAX; q++) M[q] = 0.0;
    constexpr int NFD = MODE == 2 ? 15 : (MODE == 1 ? 12 : 6);
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < n3; c += gridDim.x * blockDim.x) {
        const int k = c % N, j = (c / N) % N, i = c / ((uint32_t)N * N);
        uint32_t np[3] = {sidx3(N, wpd(i, N), j, k), sidx3(N, i, wpd(j, N), k), sidx3(N, i, j, wpd(k, N))};
        uint32_t nm[3] = {sidx3(N, wmd(i, N), j, k), sidx3(N, i, wmd(j, N), k), sidx3(N, i, j, wmd(k, N))};
        const auto xn = [&](int f, int d, int sg) -> double { return X(f, sg > 0 ? np[d] : nm[d]); };
        double xc[15] = {}, vc[15] = {};
#pragma unroll
        for (int f = 0; f < NFD; f++) { xc[f] = X(f, c); vc[f] = V(f, c); }
        const double rho = diag_site<MODE>(kp, dp, xc, vc, xn, i, j, k, S, M);
        if (MODE == 2) {
            const double Em[3] = {V(F_TH + 0, nm[0]), V(F_TH + 1, nm[1]), V(F_TH + 2, nm[2])};
            gauss_site(kp, dp, vc + F_TH, Em, rho, i, j, k, S, M);
        }
    }
    double *out = partials + (size_t)blockIdx.x * (DS_NSUM + DM_NMAX);
    block_reduce<DS_NSUM, false>(S, out);
    block_reduce<DM_NMAX, true>(M, out + DS_NSUM);
}

// pass 2: r_core second moment and the argmax-s voxel (lowest index with
// s >= smax(1 - 1e-12), C:2116-2168)
__global__ void __launch_bounds__(DIAG_BLOCK)
k_diag2(KParams kp, const double *__restrict__ xo, double thr, double cx, double cy, double cz,
        double *partials, unsigned long long *argmin) {
    const int N = kp.N;
    const uint32_t fs = kp.fs, n3 = kp.n3;
    auto X = [&](int f, uint32_t c) -> double { return xo[(size_t)f * fs + c]; };
    double v[1] = {0.0};
    unsigned long long best = ~0ull;
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < n3; c += gridDim.x * blockDim.x) {
        const int k = c % N, j = (c / N) % N, i = c / ((uint32_t)N * N);
        double x = -kp.L + i * kp.dx, y = -kp.L + j * kp.dx, z = -kp.L + k * kp.dx;
        double rho2 = 0, s = 1.0;
        for (int a = 0; a < 3; a++) {
            double u = X(F_U + a, c), w = X(F_V + a, c);
            double s2a = u * u + w * w;
            rho2 += u * u + w * w;
            s *= s2a;
        }
        v[0] += ((x - cx) * (x - cx) + (y - cy) * (y - cy) + (z - cz) * (z - cz)) * rho2;
        if (s >= thr && c < best) best = c;
    }
    for (int o = 32; o > 0; o >>= 1) {
        unsigned long long ob = __shfl_xor(best, o, 64);
        best = ob < best ? ob : best;
    }
    if ((threadIdx.x & 63) == 0 && best != ~0ull) atomicMin(argmin, best);
    block_reduce<1, false>(v, partials + (size_t)blockIdx.x);
}

// Gauss statistics of the seam sites of a fused-diagnostics launch: the sites
// whose backward neighbour (x - e_d) another block of that launch emitted — the
// first plane of each DI-plane segment, the first row of each 4-row tile, the
// first column of each 64-column tile — from the emitted E and the rho the
// launch kept for them (compute_gauss, C:2186-2243); every other site's
// statistics were accumulated by the launch itself. Seam sites are enumerated
// compactly: rows j = 0 mod 4 (every i, k), then columns k = 0 mod 64 of the
// other rows, then planes i = 0 mod DI of the rest. The row's other slots are
// zero, so it reduces with the step's partials.
__device__ inline int nth_nonmultiple(uint32_t q, int m) { return (int)((q / (m - 1)) * m + q % (m - 1) + 1); }

__global__ void __launch_bounds__(DIAG_BLOCK)
k_gauss_seam(KParams kp, DiagParams dp, int DI, const double *__restrict__ vo, const double *__restrict__ rho,
             double *partials) {
    const int N = kp.N;
    const uint32_t fs = kp.fs, uN = (uint32_t)N;
    const uint32_t nj4 = (uN + 3) / 4, nk64 = (uN + 63) / 64, niD = (uN + DI - 1) / DI;
    const uint32_t n1 = uN * nj4 * uN, n2 = uN * (uN - nj4) * nk64, n3s = niD * (uN - nj4) * (uN - nk64);
    double S[DS_NSUM], M[DM_NMAX];
#pragma unroll
    for (int q = 0; q < DS_NSUM; q++) S[q] = 0.0;
#pragma unroll
    for (int q = 0; q < DM_NMAX; q++) M[q] = 0.0;
    const double *E0 = vo + (size_t)F_TH * fs, *E1 = E0 + fs, *E2 = E1 + fs;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n1 + n2 + n3s; t += gridDim.x * blockDim.x) {
        int i, j, k;
        if (t < n1) {
            k = t % uN; j = 4 * (int)((t / uN) % nj4); i = t / (uN * nj4);
        } else if (t < n1 + n2) {
            const uint32_t q = t - n1;
            k = 64 * (int)(q % nk64); j = nth_nonmultiple((q / nk64) % (uN - nj4), 4); i = q / (nk64 * (uN - nj4));
        } else {
            const uint32_t q = t - n1 - n2;
            k = nth_nonmultiple(q % (uN - nk64), 64); j = nth_nonmultiple((q / (uN - nk64)) % (uN - nj4), 4);
            i = DI * (int)(q / ((uN - nk64) * (uN - nj4)));
        }
        const uint32_t c = sidx3(N, i, j, k);
        const double E[3] = {E0[c], E1[c], E2[c]};
        const double Em[3] = {E0[sidx3(N, wmd(i, N), j, k)], E1[sidx3(N, i, wmd(j, N), k)],
                              E2[sidx3(N, i, j, wmd(k, N))]};
        gauss_site(kp, dp, E, Em, rho[c], i, j, k, S, M);
    }
    double *out = partials + (size_t)blockIdx.x * (DS_NSUM + DM_NMAX);
    block_reduce<DS_NSUM, false>(S, out);
    block_reduce<DM_NMAX, true>(M, out + DS_NSUM);
}

// Wide form of k_reduce for many partial rows: block q reduces slot q, thread t
// rows t, t + 256, ... in order, then a fixed butterfly; deterministic. The
// maxima are of non-negative values (0 is their identity).
__global__ void __launch_bounds__(256)
k_reduce_wide(const double *partials, int nblocks, int nsum, int nmax, double *out) {
    __shared__ double red[4];
    const int q = blockIdx.x, stride = nsum + nmax, t = threadIdx.x;
    const bool isum = q < nsum;
    double x = 0.0;
    for (int b = t; b < nblocks; b += 256) {
        const double y = partials[(size_t)b * stride + q];
        x = isum ? x + y : fmax(x, y);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double y = __shfl_xor(x, o, 64);
        x = isum ? x + y : fmax(x, y);
    }
    if ((t & 63) == 0) red[t >> 6] = x;
    __syncthreads();
    if (t == 0) {
        double v = red[0];
        for (int w = 1; w < 4; w++) v = isum ? v + red[w] : fmax(v, red[w]);
        out[q] = v;
    }
}

// sums (first nsum) and maxima (next nmax) of the per-block partials, in block order
__global__ void k_reduce(const double *partials, int nblocks, int nsum, int nmax, double *out) {
    int q = threadIdx.x;
    if (q >= nsum + nmax) return;
    const int stride = nsum + nmax;
    double x = partials[q];
    for (int b = 1; b < nblocks; b++) {
        double y = partials[(size_t)b * stride + q];
        x = q < nsum ? x + y : fmax(x, y);
    }
    out[q] = x;
}

int launch_diag_pass1(int mode, const KParams &kp, const DiagParams &dp, const double *xo,
                      const double *vo, double *partials, int nblocks, hipStream_t s) {
    if (mode == 2) hipLaunchKernelGGL(k_diag1<2>, dim3(nblocks), dim3(DIAG_BLOCK), 0, s, kp, dp, xo, vo, partials);
    else if (mode == 1) hipLaunchKernelGGL(k_diag1<1>, dim3(nblocks), dim3(DIAG_BLOCK), 0, s, kp, dp, xo, vo, partials);
    else hipLaunchKernelGGL(k_diag1<0>, dim3(nblocks), dim3(DIAG_BLOCK), 0, s, kp, dp, xo, vo, partials);
    return (int)hipGetLastError();
}

int launch_diag_pass2(const KParams &kp, const double *xo, double thr, double cx, double cy,
                      double cz, double *partials, unsigned long long *argmin, int nblocks,
                      hipStream_t s) {
    hipLaunchKernelGGL(k_diag2, dim3(nblocks), dim3(DIAG_BLOCK), 0, s, kp, xo, thr, cx, cy, cz, partials, argmin);
    return (int)hipGetLastError();
}

int launch_gauss_seam(const KParams &kp, const DiagParams &dp, int DI, const double *vo, const double *rho,
                      double *partials, int nblocks, hipStream_t s) {
    if (DI <= 0) DI = march_di(kp.N, 1);
    hipLaunchKernelGGL(k_gauss_seam, dim3(nblocks), dim3(DIAG_BLOCK), 0, s, kp, dp, DI, vo, rho, partials);
    return (int)hipGetLastError();
}

int launch_reduce_wide(const double *partials, int nblocks, int nsum, int nmax, double *out, hipStream_t s) {
    hipLaunchKernelGGL(k_reduce_wide, dim3(nsum + nmax), dim3(256), 0, s, partials, nblocks, nsum, nmax, out);
    return (int)hipGetLastError();
}

int launch_reduce(const double *partials, int nblocks, int nsum, int nmax, double *out, hipStream_t s) {
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(64), 0, s, partials, nblocks, nsum, nmax, out);
    return (int)hipGetLastError();
}

// ---------------------------------------------------------------------------
// Gauss projection building blocks (init_gauss_project, C:2284-2400)
// ---------------------------------------------------------------------------

__global__ void k_gauss_residual(KParams kp, const double *xo, const double *vo, double *b) {
    const int N = kp.N;
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= kp.n3) return;
    const uint32_t fs = kp.fs;
    const int k = c % N, j = (c / N) % N, i = c / ((uint32_t)N * N);
    uint32_t nm[3] = {sidx3(N, wmd(i, N), j, k), sidx3(N, i, wmd(j, N), k), sidx3(N, i, j, wmd(k, N))};
    auto X = [&](int f, uint32_t q) -> double { return xo[(size_t)f * fs + q]; };
    auto V = [&](int f, uint32_t q) -> double { return vo[(size_t)f * fs + q]; };
    double divE = (V(F_TH, c) - V(F_TH, nm[0]) + V(F_TH + 1, c) - V(F_TH + 1, nm[1])
                 + V(F_TH + 2, c) - V(F_TH + 2, nm[2])) / kp.dx;
    double rho = 0;
    for (int a = 0; a < 3; a++)
        rho += X(F_U + a, c) * V(F_V + a, c) - X(F_V + a, c) * V(F_U + a, c)
             + X(F_TU + a, c) * V(F_TV + a, c) - X(F_TV + a, c) * V(F_TU + a, c);
    b[c] = divE - kp.GG * rho;
}

__global__ void k_neg_lap(KParams kp, const double *p, double *ap) {
    const int N = kp.N;
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= kp.n3) return;
    const int k = c % N, j = (c / N) % N, i = c / ((uint32_t)N * N);
    double lap = (p[sidx3(N, wpd(i, N), j, k)] + p[sidx3(N, wmd(i, N), j, k)] + p[sidx3(N, i, wpd(j, N), k)]
                + p[sidx3(N, i, wmd(j, N), k)] + p[sidx3(N, i, j, wpd(k, N))] + p[sidx3(N, i, j, wmd(k, N))]
                - 6.0 * p[c]) * kp.h2;
    ap[c] = -lap;
}

__global__ void k_add_grad(KParams kp, const double *chi, double *E) {
    const int N = kp.N;
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= kp.n3) return;
    const int k = c % N, j = (c / N) % N, i = c / ((uint32_t)N * N);
    uint32_t np[3] = {sidx3(N, wpd(i, N), j, k), sidx3(N, i, wpd(j, N), k), sidx3(N, i, j, wpd(k, N))};
    for (int d = 0; d < 3; d++) E[(size_t)d * kp.fs + c] += (chi[np[d]] - chi[c]) / kp.dx;
}

// partials[b] = sum over the block's sites of a*b
__global__ void __launch_bounds__(DIAG_BLOCK) k_dot(const double *a, const double *b, uint32_t n, double *partials) {
    double v[1] = {0.0};
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) v[0] += a[c] * b[c];
    block_reduce<1, false>(v, partials + blockIdx.x);
}

// x += alpha p; r -= alpha Ap; partials = {r.r, sum x, max|r|}
__global__ void __launch_bounds__(DIAG_BLOCK)
k_axpy_cg(double alpha, const double *p, const double *ap, double *x, double *r, uint32_t n, double *partials) {
    double sv[2] = {0.0, 0.0}, mv[1] = {0.0};
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
        x[c] += alpha * p[c];
        r[c] -= alpha * ap[c];
        sv[0] += r[c] * r[c];
        sv[1] += x[c];
        mv[0] = fmax(mv[0], fabs(r[c]));
    }
    block_reduce<2, false>(sv, partials + (size_t)blockIdx.x * 3);
    block_reduce<1, true>(mv, partials + (size_t)blockIdx.x * 3 + 2);
}

__global__ void k_cg_update(double xbar, double beta, const double *r, double *x, double *p, uint32_t n) {
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    x[c] -= xbar;
    p[c] = r[c] + beta * p[c];
}

// b -= mean; partials = {max|b|}
__global__ void __launch_bounds__(DIAG_BLOCK) k_sub_mean_max(double *b, double mean, uint32_t n, double *partials) {
    double mv[1] = {0.0};
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
        b[c] -= mean;
        mv[0] = fmax(mv[0], fabs(b[c]));
    }
    block_reduce<1, true>(mv, partials + blockIdx.x);
}

__global__ void k_scale(double *x, double f, uint32_t n) {
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n) x[c] *= f;
}

static inline dim3 g1(uint32_t n) { return dim3((n + 255) / 256); }

int launch_gauss_residual(const KParams &kp, const double *xo, const double *vo, double *b, hipStream_t s) {
    hipLaunchKernelGGL(k_gauss_residual, g1(kp.n3), dim3(256), 0, s, kp, xo, vo, b);
    return (int)hipGetLastError();
}
int launch_neg_lap(const KParams &kp, const double *p, double *ap, hipStream_t s) {
    hipLaunchKernelGGL(k_neg_lap, g1(kp.n3), dim3(256), 0, s, kp, p, ap);
    return (int)hipGetLastError();
}
int launch_add_grad(const KParams &kp, const double *chi, double *E, hipStream_t s) {
    hipLaunchKernelGGL(k_add_grad, g1(kp.n3), dim3(256), 0, s, kp, chi, E);
    return (int)hipGetLastError();
}
int launch_dot_partials(const double *a, const double *b, uint32_t n, double *partials, int nblocks, hipStream_t s) {
    hipLaunchKernelGGL(k_dot, dim3(nblocks), dim3(DIAG_BLOCK), 0, s, a, b, n, partials);
    return (int)hipGetLastError();
}
int launch_axpy_cg(double alpha, const double *p, const double *ap, double *x, double *r, uint32_t n,
                   double *partials, int nblocks, hipStream_t s) {
    hipLaunchKernelGGL(k_axpy_cg, dim3(nblocks), dim3(DIAG_BLOCK), 0, s, alpha, p, ap, x, r, n, partials);
    return (int)hipGetLastError();
}
int launch_cg_update(double xbar, double beta, const double *r, double *x, double *p, uint32_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_cg_update, g1(n), dim3(256), 0, s, xbar, beta, r, x, p, n);
    return (int)hipGetLastError();
}
int launch_sub_mean_max(double *b, double mean, uint32_t n, double *partials, int nblocks, hipStream_t s) {
    hipLaunchKernelGGL(k_sub_mean_max, dim3(nblocks), dim3(DIAG_BLOCK), 0, s, b, mean, n, partials);
    return (int)hipGetLastError();
}
int launch_scale(double *x, double f, uint32_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_scale, g1(n), dim3(256), 0, s, x, f, n);
    return (int)hipGetLastError();
}

}  // namespace scp
