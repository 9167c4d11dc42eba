// scp_diag.hip — FP64 diagnostics and Gauss-law kernels.
//
// Diagnostics (reference scp_sim.c:1722-2243) are one fused pass over the
// emitted state: every energy term, Noether charges, P_int, theta_rms, the
// Gauss residual statistics, E_em and Q_flux accumulate per thread, reduce per
// block with a fixed butterfly, and a single-block kernel sums the block
// partials in block order. The result is bitwise reproducible run to run
// (the reference's OpenMP/atomicAdd sums are not; they agree to a few ulps).
#include "scp_kernels.h"

#include <hip/hip_runtime.h>

namespace scp {

__device__ inline uint32_t sidx3(int N, int i, int j, int k) {
    return ((uint32_t)i * (uint32_t)N + (uint32_t)j) * (uint32_t)N + (uint32_t)k;
}
__device__ inline int wpd(int a, int N) { return a + 1 == N ? 0 : a + 1; }
__device__ inline int wmd(int a, int N) { return a == 0 ? N - 1 : a - 1; }

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
__global__ void __launch_bounds__(DIAG_BLOCK)
k_diag1(KParams kp, DiagParams dp, const double *__restrict__ xo, const double *__restrict__ vo,
        double *partials) {
    const int N = kp.N;
    const uint32_t fs = kp.fs, n3 = kp.n3;
    const double dx = kp.dx, dV = dp.dV, h1 = kp.h1, L = kp.L;
    auto X = [&](int f, uint32_t c) -> double { return xo[(size_t)f * fs + c]; };
    auto V = [&](int f, uint32_t c) -> double { return vo[(size_t)f * fs + c]; };
    double S[DS_NSUM], M[DM_NMAX];
#pragma unroll
    for (int q = 0; q < DS_NSUM; q++) S[q] = 0.0;
#pragma unroll
    for (int q = 0; q < DM_NMAX; q++) M[q] = 0.0;
    const int C1[3] = {1, 2, 0}, C2[3] = {2, 0, 1};
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < n3; c += gridDim.x * blockDim.x) {
        const int k = c % N, j = (c / N) % N, i = c / ((uint32_t)N * N);
        uint32_t np[3] = {sidx3(N, wpd(i, N), j, k), sidx3(N, i, wpd(j, N), k), sidx3(N, i, j, wpd(k, N))};
        uint32_t nm[3] = {sidx3(N, wmd(i, N), j, k), sidx3(N, i, wmd(j, N), k), sidx3(N, i, j, wmd(k, N))};
        const double x = -L + i * dx, y = -L + j * dx, z = -L + k * dx;
        if constexpr (MODE == 0) {              // compute_energy, C:1734-1796 (+ C:1803-1816)
            const int TU = 3;
            double p0 = X(0, c), p1 = X(1, c), p2 = X(2, c);
            double sig = p0 * p0 + p1 * p1 + p2 * p2;
            double me2 = (kp.rmode == 1) ? kp.IA / (1.0 + kp.IB * sig) : kp.MASS2;
            double keff = (kp.rmode == 3) ? kp.KAPPA / (1.0 + kp.KG * sig) : kp.KAPPA;
            for (int a = 0; a < 3; a++) {
                S[DS_EPK] += 0.5 * V(a, c) * V(a, c) * dV;
                S[DS_ETK] += 0.5 * V(TU + a, c) * V(TU + a, c) * dV;
                double gx = (X(a, np[0]) - X(a, nm[0])) * h1, gy = (X(a, np[1]) - X(a, nm[1])) * h1;
                double gz = (X(a, np[2]) - X(a, nm[2])) * h1;
                S[DS_EG] += 0.5 * (gx * gx + gy * gy + gz * gz) * dV;
                S[DS_EM] += 0.5 * me2 * X(a, c) * X(a, c) * dV;
                double tx = (X(TU + a, np[0]) - X(TU + a, nm[0])) * h1, ty = (X(TU + a, np[1]) - X(TU + a, nm[1])) * h1;
                double tz = (X(TU + a, np[2]) - X(TU + a, nm[2])) * h1;
                S[DS_ETG] += 0.5 * (tx * tx + ty * ty + tz * tz) * dV;
                S[DS_ETM] += 0.5 * kp.MT2 * X(TU + a, c) * X(TU + a, c) * dV;
                M[DM_PHI] = fmax(M[DM_PHI], fabs(X(a, c)));
                S[DS_TH2] += X(TU + a, c) * X(TU + a, c);
            }
            double P = p0 * p1 * p2, P2 = P * P;
            if (kp.rmode == 3) {
                double D = 1.0 + kp.KG * sig + kp.KAPPA * P2;
                S[DS_EP] += (kp.MU / 2.0) * P2 * (1.0 + kp.KG * sig) / D * dV;
            } else {
                S[DS_EP] += (kp.MU / 2.0) * P2 / (1.0 + keff * P2) * dV;
            }
            M[DM_P] = fmax(M[DM_P], fabs(P));
            if (kp.ACS != 0 || kp.BH != 0 || kp.KH != 0) {
                double cx[3];
                cx[0] = (X(2, np[1]) - X(2, nm[1]) - X(1, np[2]) + X(1, nm[2])) * h1;
                cx[1] = (X(0, np[2]) - X(0, nm[2]) - X(2, np[0]) + X(2, nm[0])) * h1;
                cx[2] = (X(1, np[0]) - X(1, nm[0]) - X(0, np[1]) + X(0, nm[1])) * h1;
                if (kp.ACS != 0) {
                    double m0 = cx[0] * 0.5 - X(3, c), m1 = cx[1] * 0.5 - X(4, c), m2 = cx[2] * 0.5 - X(5, c);
                    S[DS_EP] += kp.ACS * (m0 * m0 + m1 * m1 + m2 * m2) * dV;
                }
                if (kp.BH != 0) {
                    double csq = cx[0] * cx[0] + cx[1] * cx[1] + cx[2] * cx[2];
                    double T2 = X(3, c) * X(3, c) + X(4, c) * X(4, c) + X(5, c) * X(5, c);
                    S[DS_EP] += 0.5 * kp.BH * T2 * csq * dV;
                }
                if (kp.KH != 0) {
                    double hel = p0 * cx[0] + p1 * cx[1] + p2 * cx[2];
                    S[DS_EP] += kp.KH * P2 * hel * dV;
                }
            }
            double ct[3];
            ct[0] = (X(TU + 2, np[1]) - X(TU + 2, nm[1]) - X(TU + 1, np[2]) + X(TU + 1, nm[2])) * h1;
            ct[1] = (X(TU + 0, np[2]) - X(TU + 0, nm[2]) - X(TU + 2, np[0]) + X(TU + 2, nm[0])) * h1;
            ct[2] = (X(TU + 1, np[0]) - X(TU + 1, nm[0]) - X(TU + 0, np[1]) + X(TU + 0, nm[1])) * h1;
            for (int a = 0; a < 3; a++) S[DS_EC] -= kp.ETA * X(a, c) * ct[a] * dV;
            S[DS_PINT] += fabs(p0 * p1 * p2) * dV;
        } else {
            const bool gauged = (MODE == 2) && dp.gauged;
            double fu[6], fv[6];
            for (int a = 0; a < 3; a++) {
                fu[a] = X(F_U + a, c); fv[a] = X(F_V + a, c);
                fu[3 + a] = X(F_TU + a, c); fv[3 + a] = X(F_TV + a, c);
            }
            double DU[3][6], DV[3][6];
            for (int d = 0; d < 3; d++) {
                double cP = 1, sP = 0, cM = 1, sM = 0;
                if (gauged) { sincos(X(F_TH + d, c), &sP, &cP); sincos(X(F_TH + d, nm[d]), &sM, &cM); }
                for (int a = 0; a < 3; a++)
                    for (int h = 0; h < 2; h++) {
                        const int fu_ = h ? F_TU + a : F_U + a, fv_ = h ? F_TV + a : F_V + a;
                        double up = X(fu_, np[d]), vp = X(fv_, np[d]), um = X(fu_, nm[d]), vm = X(fv_, nm[d]);
                        if (gauged) {
                            DU[d][h * 3 + a] = ((cP * up - sP * vp) - (cM * um + sM * vm)) * h1;
                            DV[d][h * 3 + a] = ((cP * vp + sP * up) - (cM * vm - sM * um)) * h1;
                        } else {
                            DU[d][h * 3 + a] = (up - um) * h1;
                            DV[d][h * 3 + a] = (vp - vm) * h1;
                        }
                    }
            }
            double s2[3], rho2 = 0, s = 1.0, qpv = 0, qtv = 0;
            for (int a = 0; a < 3; a++) {
                s2[a] = fu[a] * fu[a] + fv[a] * fv[a];
                double ud = V(F_U + a, c), vd = V(F_V + a, c), tud = V(F_TU + a, c), tvd = V(F_TV + a, c);
                S[DS_EPK] += 0.5 * (ud * ud + vd * vd) * dV;
                S[DS_ETK] += 0.5 * (tud * tud + tvd * tvd) * dV;
                S[DS_EG] += 0.5 * (DU[0][a] * DU[0][a] + DU[1][a] * DU[1][a] + DU[2][a] * DU[2][a]
                                 + DV[0][a] * DV[0][a] + DV[1][a] * DV[1][a] + DV[2][a] * DV[2][a]) * dV;
                if (gauged) S[DS_EM] += 0.5 * kp.MASS2 * s2[a] * dV;
                else S[DS_EM] += 0.5 * kp.MASS2 * (fu[a] * fu[a] + fv[a] * fv[a]) * dV;
                S[DS_ETG] += 0.5 * (DU[0][3 + a] * DU[0][3 + a] + DU[1][3 + a] * DU[1][3 + a] + DU[2][3 + a] * DU[2][3 + a]
                                  + DV[0][3 + a] * DV[0][3 + a] + DV[1][3 + a] * DV[1][3 + a] + DV[2][3 + a] * DV[2][3 + a]) * dV;
                S[DS_ETM] += 0.5 * kp.MT2 * (fu[3 + a] * fu[3 + a] + fv[3 + a] * fv[3 + a]) * dV;
                M[DM_PHI] = fmax(M[DM_PHI], sqrt(s2[a]));
                double reDxT, imDxT;
                if (gauged) {
                    reDxT = DU[C1[a]][3 + C2[a]] - DU[C2[a]][3 + C1[a]];
                    imDxT = DV[C1[a]][3 + C2[a]] - DV[C2[a]][3 + C1[a]];
                } else {   // ungauged curls read the raw neighbours (C:1868-1869)
                    const int b1 = C1[a], b2 = C2[a];
                    reDxT = (X(F_TU + b2, np[b1]) - X(F_TU + b2, nm[b1]) - X(F_TU + b1, np[b2]) + X(F_TU + b1, nm[b2])) * h1;
                    imDxT = (X(F_TV + b2, np[b1]) - X(F_TV + b2, nm[b1]) - X(F_TV + b1, np[b2]) + X(F_TV + b1, nm[b2])) * h1;
                }
                S[DS_EC] -= kp.ETA * (fu[a] * reDxT + fv[a] * imDxT) * dV;
                S[DS_TH2] += fu[3 + a] * fu[3 + a] + fv[3 + a] * fv[3 + a];
                // charges (C:2091-2103)
                double q = fu[a] * vd - fv[a] * ud;
                qpv += q;
                S[DS_QP0 + a] += q;
                qtv += fu[3 + a] * tvd - fv[3 + a] * tud;
                rho2 += s2[a];
                s *= s2[a];
            }
            double sp = s2[0] * s2[1] * s2[2];
            S[DS_EP] += (kp.MU / 2.0) * sp / (1.0 + kp.KAPPA * sp) * dV;
            M[DM_P] = fmax(M[DM_P], sqrt(sp));
            S[DS_PINT] += sqrt(sp) * dV;
            S[DS_QP] += qpv; S[DS_QT] += qtv;
            if (x * x + y * y + z * z <= dp.R2core) S[DS_QC] += qpv + qtv;
            M[DM_S] = fmax(M[DM_S], s);
            S[DS_M0] += rho2; S[DS_CX] += x * rho2; S[DS_CY] += y * rho2; S[DS_CZ] += z * rho2;
            if (MODE == 2) {
                double E0 = V(F_TH + 0, c), E1 = V(F_TH + 1, c), E2 = V(F_TH + 2, c);
                double ee = 0.5 * (E0 * E0 + E1 * E1 + E2 * E2);
                double bb = 0.0;
                if (gauged) {
                    const int PA[3] = {0, 1, 2}, PB[3] = {1, 2, 0};
                    for (int p = 0; p < 3; p++) {
                        int a = PA[p], b = PB[p];
                        double ang = X(F_TH + a, c) + X(F_TH + b, np[a]) - X(F_TH + a, np[b]) - X(F_TH + b, c);
                        bb += 1.0 - cos(ang);
                    }
                    S[DS_EEM_E] += (ee + dp.BPRE * bb) * dV;
                }
                // Gauss residual (C:2176-2181, 2210-2219)
                double divE = (E0 - V(F_TH + 0, nm[0]) + E1 - V(F_TH + 1, nm[1]) + E2 - V(F_TH + 2, nm[2])) / dx;
                double rho = 0;
                for (int a = 0; a < 3; a++)
                    rho += fu[a] * V(F_V + a, c) - fv[a] * V(F_U + a, c)
                         + fu[3 + a] * V(F_TV + a, c) - fv[3 + a] * V(F_TU + a, c);
                double G = divE - kp.GG * rho;
                if (x * x + y * y + z * z < dp.Rint2) {
                    double dG = fabs(G - dp.G_offset);
                    M[DM_GAUSS] = fmax(M[DM_GAUSS], dG);
                    S[DS_GL2] += (G - dp.G_offset) * (G - dp.G_offset);
                    S[DS_NOMEGA] += 1.0;
                }
                double eg = ee;
                if (kp.GG != 0) eg += bb / (kp.GG * kp.GG * dx * dx * dx * dx);
                S[DS_EEM_G] += eg * dV;
                if (fabs(x) <= dp.Rcube && fabs(y) <= dp.Rcube && fabs(z) <= dp.Rcube) {
                    S[DS_FSUM] += divE * dV;
                    S[DS_NC] += 1.0;
                }
            }
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
