// scp_diag_site.h — per-site diagnostics of the three modes,
// shared by the stand-alone diagnostics pass (scp_diag.hip, neighbours from the
// emitted FP64 state) and the diagnostics fused into the step's emitting
// launch (scp_march.h, neighbours from the staged planes).
//
// Energy terms (compute_energy_complex, scp_sim.c:1823-1880, and
// compute_energy_complex_gauge, C:1886-2036), Noether charges and the core
// moments (compute_charges, C:2074-2169), P_int and theta_rms (C:2039-2059) of
// one site, accumulated into S[] and M[] (the DS_* / DM_* slots). Everything
// but the Gauss residual, which needs E at the backward neighbours: the
// function returns the site's charge density rho (C:2176-2181) for it.
#pragma once

#include "scp_physics.h"

#include <hip/hip_runtime.h>

namespace scp {

// Real-mode terms of one site (compute_energy, C:1734-1796, with the alpha_cs,
// beta_h, kappa_h and coupling terms of C:1803-1816; P_int and theta_rms of
// C:2039-2059): xc/vc the six fields and velocities (phi 0..2, theta 3..5), xn
// as below.
template <class XN>
__device__ inline void diag_site_real(const KParams &kp, const DiagParams &dp, const double *xc, const double *vc,
                                      const XN &xn, double *S, double *M) {
#pragma clang fp contract(off)
    const double dV = dp.dV, h1 = kp.h1;
    const int TU = 3;
    double p0 = xc[0], p1 = xc[1], p2 = xc[2];
    double sig = p0 * p0 + p1 * p1 + p2 * p2;
    double me2 = (kp.rmode == 1) ? kp.IA / (1.0 + kp.IB * sig) : kp.MASS2;
    double keff = (kp.rmode == 3) ? kp.KAPPA / (1.0 + kp.KG * sig) : kp.KAPPA;
    for (int a = 0; a < 3; a++) {
        S[DS_EPK] += 0.5 * vc[a] * vc[a] * dV;
        S[DS_ETK] += 0.5 * vc[TU + a] * vc[TU + a] * dV;
        double gx = (xn(a, 0, 1) - xn(a, 0, -1)) * h1, gy = (xn(a, 1, 1) - xn(a, 1, -1)) * h1;
        double gz = (xn(a, 2, 1) - xn(a, 2, -1)) * h1;
        S[DS_EG] += 0.5 * (gx * gx + gy * gy + gz * gz) * dV;
        S[DS_EM] += 0.5 * me2 * xc[a] * xc[a] * dV;
        double tx = (xn(TU + a, 0, 1) - xn(TU + a, 0, -1)) * h1, ty = (xn(TU + a, 1, 1) - xn(TU + a, 1, -1)) * h1;
        double tz = (xn(TU + a, 2, 1) - xn(TU + a, 2, -1)) * h1;
        S[DS_ETG] += 0.5 * (tx * tx + ty * ty + tz * tz) * dV;
        S[DS_ETM] += 0.5 * kp.MT2 * xc[TU + a] * xc[TU + a] * dV;
        M[DM_PHI] = fmax(M[DM_PHI], fabs(xc[a]));
        S[DS_TH2] += xc[TU + a] * xc[TU + a];
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
        cx[0] = (xn(2, 1, 1) - xn(2, 1, -1) - xn(1, 2, 1) + xn(1, 2, -1)) * h1;
        cx[1] = (xn(0, 2, 1) - xn(0, 2, -1) - xn(2, 0, 1) + xn(2, 0, -1)) * h1;
        cx[2] = (xn(1, 0, 1) - xn(1, 0, -1) - xn(0, 1, 1) + xn(0, 1, -1)) * h1;
        if (kp.ACS != 0) {
            double m0 = cx[0] * 0.5 - xc[3], m1 = cx[1] * 0.5 - xc[4], m2 = cx[2] * 0.5 - xc[5];
            S[DS_EP] += kp.ACS * (m0 * m0 + m1 * m1 + m2 * m2) * dV;
        }
        if (kp.BH != 0) {
            double csq = cx[0] * cx[0] + cx[1] * cx[1] + cx[2] * cx[2];
            double T2 = xc[3] * xc[3] + xc[4] * xc[4] + xc[5] * xc[5];
            S[DS_EP] += 0.5 * kp.BH * T2 * csq * dV;
        }
        if (kp.KH != 0) {
            double hel = p0 * cx[0] + p1 * cx[1] + p2 * cx[2];
            S[DS_EP] += kp.KH * P2 * hel * dV;
        }
    }
    double ct[3];
    ct[0] = (xn(TU + 2, 1, 1) - xn(TU + 2, 1, -1) - xn(TU + 1, 2, 1) + xn(TU + 1, 2, -1)) * h1;
    ct[1] = (xn(TU + 0, 2, 1) - xn(TU + 0, 2, -1) - xn(TU + 2, 0, 1) + xn(TU + 2, 0, -1)) * h1;
    ct[2] = (xn(TU + 1, 0, 1) - xn(TU + 1, 0, -1) - xn(TU + 0, 1, 1) + xn(TU + 0, 1, -1)) * h1;
    for (int a = 0; a < 3; a++) S[DS_EC] -= kp.ETA * xc[a] * ct[a] * dV;
    S[DS_PINT] += fabs(p0 * p1 * p2) * dV;
}

// xc, vc: the site's fields and velocities (device field order, FP64);
// xn(f, d, sg): field f at the neighbour one step along d (0 = i, 1 = j,
// 2 = k) in direction sg = +1 / -1, as a double. Returns the Gauss-law charge
// density (gauged mode; 0 otherwise).
template <int MODE, class XN>
__device__ inline double diag_site(const KParams &kp, const DiagParams &dp, const double *xc, const double *vc,
                                   const XN &xn, int i, int j, int k, double *S, double *M) {
#pragma clang fp contract(off)
    if constexpr (MODE == 0) {
        diag_site_real(kp, dp, xc, vc, xn, S, M);
        return 0.0;
    }
    const double dx = kp.dx, dV = dp.dV, h1 = kp.h1, L = kp.L;
    const double x = -L + i * dx, y = -L + j * dx, z = -L + k * dx;
    const bool gauged = (MODE == 2) && dp.gauged;
    const int C1[3] = {1, 2, 0}, C2[3] = {2, 0, 1};
    double fu[6], fv[6];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        fu[a] = xc[F_U + a]; fv[a] = xc[F_V + a];
        fu[3 + a] = xc[F_TU + a]; fv[3 + a] = xc[F_TV + a];
    }
    // covariant (gauged) or plain central differences, C:1838-1849, 1915-1931,
    // formed per field as consumed (only the six Theta differences the
    // covariant curls need are kept); every sum keeps the reference's order
    double cP[3], sP[3], cM[3], sM[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        cP[d] = 1; sP[d] = 0; cM[d] = 1; sM[d] = 0;
        if (gauged) {   // this build's FP64 sincos, as in the step (scp_physics.h)
            const SinCos64 p = sincos64(xc[F_TH + d]), m = sincos64(xn(F_TH + d, d, -1));
            sP[d] = p.s; cP[d] = p.c; sM[d] = m.s; cM[d] = m.c;
        }
    }
    // D of field pair (fu_, fv_) along d: (re, im)
    auto D = [&](int fu_, int fv_, int d, double &re, double &im) {
        const double up = xn(fu_, d, 1), vp = xn(fv_, d, 1), um = xn(fu_, d, -1), vm = xn(fv_, d, -1);
        if (gauged) {
            re = ((cP[d] * up - sP[d] * vp) - (cM[d] * um + sM[d] * vm)) * h1;
            im = ((cP[d] * vp + sP[d] * up) - (cM[d] * vm - sM[d] * um)) * h1;
        } else {
            re = (up - um) * h1;
            im = (vp - vm) * h1;
        }
    };
#pragma unroll
    for (int a = 0; a < 3; a++) {          // gradient energy of Phi_a
        double u0, v0, u1, v1, u2, v2;
        D(F_U + a, F_V + a, 0, u0, v0);
        D(F_U + a, F_V + a, 1, u1, v1);
        D(F_U + a, F_V + a, 2, u2, v2);
        S[DS_EG] += 0.5 * (u0 * u0 + u1 * u1 + u2 * u2 + v0 * v0 + v1 * v1 + v2 * v2) * dV;
    }
    double TU[3][3], TV[3][3];             // Theta differences [d][b], d != b (curls)
#pragma unroll
    for (int b = 0; b < 3; b++) {          // gradient energy of Theta_b
        double u[3], v[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            D(F_TU + b, F_TV + b, d, u[d], v[d]);
            TU[d][b] = u[d];
            TV[d][b] = v[d];
        }
        S[DS_ETG] += 0.5 * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2] + v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) * dV;
    }
    double s2[3], rho2 = 0, s = 1.0, qpv = 0, qtv = 0;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        s2[a] = fu[a] * fu[a] + fv[a] * fv[a];
        const double ud = vc[F_U + a], vd = vc[F_V + a], tud = vc[F_TU + a], tvd = vc[F_TV + a];
        S[DS_EPK] += 0.5 * (ud * ud + vd * vd) * dV;
        S[DS_ETK] += 0.5 * (tud * tud + tvd * tvd) * dV;
        if (gauged) S[DS_EM] += 0.5 * kp.MASS2 * s2[a] * dV;
        else S[DS_EM] += 0.5 * kp.MASS2 * (fu[a] * fu[a] + fv[a] * fv[a]) * dV;
        S[DS_ETM] += 0.5 * kp.MT2 * (fu[3 + a] * fu[3 + a] + fv[3 + a] * fv[3 + a]) * dV;
        M[DM_PHI] = fmax(M[DM_PHI], sqrt(s2[a]));
        double reDxT, imDxT;
        if (gauged) {
            reDxT = TU[C1[a]][C2[a]] - TU[C2[a]][C1[a]];
            imDxT = TV[C1[a]][C2[a]] - TV[C2[a]][C1[a]];
        } else {   // ungauged curls read the raw neighbours (C:1868-1869)
            const int b1 = C1[a], b2 = C2[a];
            reDxT = (xn(F_TU + b2, b1, 1) - xn(F_TU + b2, b1, -1) - xn(F_TU + b1, b2, 1) + xn(F_TU + b1, b2, -1)) * h1;
            imDxT = (xn(F_TV + b2, b1, 1) - xn(F_TV + b2, b1, -1) - xn(F_TV + b1, b2, 1) + xn(F_TV + b1, b2, -1)) * h1;
        }
        S[DS_EC] -= kp.ETA * (fu[a] * reDxT + fv[a] * imDxT) * dV;
        S[DS_TH2] += fu[3 + a] * fu[3 + a] + fv[3 + a] * fv[3 + a];
        // charges (C:2091-2103)
        const double q = fu[a] * vd - fv[a] * ud;
        qpv += q;
        S[DS_QP0 + a] += q;
        qtv += fu[3 + a] * tvd - fv[3 + a] * tud;
        rho2 += s2[a];
        s *= s2[a];
    }
    const double sp = s2[0] * s2[1] * s2[2];
    S[DS_EP] += (kp.MU / 2.0) * sp / (1.0 + kp.KAPPA * sp) * dV;
    M[DM_P] = fmax(M[DM_P], sqrt(sp));
    S[DS_PINT] += sqrt(sp) * dV;
    S[DS_QP] += qpv; S[DS_QT] += qtv;
    if (x * x + y * y + z * z <= dp.R2core) S[DS_QC] += qpv + qtv;
    M[DM_S] = fmax(M[DM_S], s);
    S[DS_M0] += rho2; S[DS_CX] += x * rho2; S[DS_CY] += y * rho2; S[DS_CZ] += z * rho2;
    double rho = 0;
    if (MODE == 2) {
        const double E0 = vc[F_TH + 0], E1 = vc[F_TH + 1], E2 = vc[F_TH + 2];
        const double ee = 0.5 * (E0 * E0 + E1 * E1 + E2 * E2);
        double bb = 0.0;
        if (gauged) {   // plaquette energy (C:1957-1968)
            const int PA[3] = {0, 1, 2}, PB[3] = {1, 2, 0};
#pragma unroll
            for (int p = 0; p < 3; p++) {
                const int a = PA[p], b = PB[p];
                const double ang = xc[F_TH + a] + xn(F_TH + b, a, 1) - xn(F_TH + a, b, 1) - xc[F_TH + b];
                bb += 1.0 - sincos64(ang).c;
            }
            S[DS_EEM_E] += (ee + dp.BPRE * bb) * dV;
        }
        double eg = ee;
        if (kp.GG != 0) eg += bb / (kp.GG * kp.GG * dx * dx * dx * dx);
        S[DS_EEM_G] += eg * dV;
        // charge density of the Gauss law (C:2176-2181)
#pragma unroll
        for (int a = 0; a < 3; a++)
            rho += fu[a] * vc[F_V + a] - fv[a] * vc[F_U + a] + fu[3 + a] * vc[F_TV + a] - fv[3 + a] * vc[F_TU + a];
    }
    return rho;
}

// Gauss-law statistics of one site (compute_gauss, C:2186-2243): G from the
// site's E, the backward neighbours' E (Em[d] = E_d(x - e_d)) and rho.
__device__ inline void gauss_site(const KParams &kp, const DiagParams &dp, const double *E, const double *Em,
                                  double rho, int i, int j, int k, double *S, double *M) {
#pragma clang fp contract(off)
    const double dx = kp.dx, dV = dp.dV, L = kp.L;
    const double x = -L + i * dx, y = -L + j * dx, z = -L + k * dx;
    const double divE = (E[0] - Em[0] + E[1] - Em[1] + E[2] - Em[2]) / dx;
    const double G = divE - kp.GG * rho;
    if (x * x + y * y + z * z < dp.Rint2) {
        const double dG = fabs(G - dp.G_offset);
        M[DM_GAUSS] = fmax(M[DM_GAUSS], dG);
        S[DS_GL2] += (G - dp.G_offset) * (G - dp.G_offset);
        S[DS_NOMEGA] += 1.0;
    }
    if (fabs(x) <= dp.Rcube && fabs(y) <= dp.Rcube && fabs(z) <= dp.Rcube) {
        S[DS_FSUM] += divE * dV;
        S[DS_NC] += 1.0;
    }
}

}  // namespace scp
