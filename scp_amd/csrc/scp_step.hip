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
#include "scp_kernels.h"

#include <hip/hip_runtime.h>

namespace scp {

#define SCP_PI 3.14159265358979323846

constexpr int TK = 64, TJ = 4;

struct Geo {
    int N, i0, i1, j, k;
    bool valid;
};

// XCD-aware block numbering: hardware block b runs on XCD b % 8; give each XCD
// a contiguous range of logical tiles so j-adjacent tiles share an L2.
__device__ inline uint32_t logical_block(uint32_t b, uint32_t nb) {
    const uint32_t nx = 8;
    uint32_t per = (nb + nx - 1) / nx;
    uint32_t xcd = b % nx, pos = b / nx;
    uint32_t l = xcd * per + pos;
    // ragged tail: fall back to identity for the few blocks past the balanced part
    return (nb % nx == 0) ? l : b;
}

__device__ inline Geo geometry(int N, int DI) {
    Geo g;
    g.N = N;
    const uint32_t ntk = (N + TK - 1) / TK, ntj = (N + TJ - 1) / TJ;
    const uint32_t nb = gridDim.x;
    uint32_t lb = logical_block(blockIdx.x, nb);
    uint32_t tk = lb % ntk; lb /= ntk;
    uint32_t tj = lb % ntj; lb /= ntj;
    uint32_t seg = lb;
    g.k = tk * TK + threadIdx.x;
    g.j = tj * TJ + threadIdx.y;
    g.i0 = seg * DI;
    g.i1 = min(N, g.i0 + DI);
    g.valid = g.k < N && g.j < N;
    return g;
}

__device__ inline uint32_t sidx(int N, int i, int j, int k) {
    return ((uint32_t)i * (uint32_t)N + (uint32_t)j) * (uint32_t)N + (uint32_t)k;
}
__device__ inline int wp(int a, int N) { return a + 1 == N ? 0 : a + 1; }
__device__ inline int wm(int a, int N) { return a == 0 ? N - 1 : a - 1; }

// ---------------------------------------------------------------------------
// storage policies
// ---------------------------------------------------------------------------

// FP64: plain doubles.
struct S64 {
    using T = double;
    using R = double;
};

// ---------------------------------------------------------------------------
// per-site physics (expression trees of scp_sim.c, kept operation for operation)
// ---------------------------------------------------------------------------

// sin of the plaquette angle th_a(y) + th_b(y+a) - th_a(y+b) - th_b(y)  (C:1100-1102)
__device__ inline double plaq(double ta_y, double tb_ya, double ta_yb, double tb_y) {
    double ang = ta_y + tb_ya - ta_yb - tb_y;
    return sin(ang);
}

template <int NP>
struct Col { double v[NP]; };

// Complex-mode accelerations at one site (compute_forces_complex, C:931-994).
// c/ip/im: own column at planes i, i+1, i-1; X(f, idx) loads neighbour values.
template <class LD>
__device__ inline void forces_complex(const KParams &kp, const double *c, const double *ip,
                                      const double *im, LD X, uint32_t jp, uint32_t jm,
                                      uint32_t kp_, uint32_t km, double *acc) {
    const double h1 = kp.h1, h2 = kp.h2;
    double s2[3];
#pragma unroll
    for (int a = 0; a < 3; a++) s2[a] = c[F_U + a] * c[F_U + a] + c[F_V + a] * c[F_V + a];
    double s = s2[0] * s2[1] * s2[2];
    double den = 1.0 + kp.KAPPA * s;
    double Vp = 0.5 * kp.MU / (den * den);
    double rest[3] = {s2[1] * s2[2], s2[0] * s2[2], s2[0] * s2[1]};
    // neighbour value of field f in direction d (+1/-1)
    auto nbv = [&](int f, int d, int sg) -> double {
        if (d == 0) return sg > 0 ? ip[f] : im[f];
        if (d == 1) return X(f, sg > 0 ? jp : jm);
        return X(f, sg > 0 ? kp_ : km);
    };
    auto lap = [&](int f) -> double {
        return (nbv(f, 0, 1) + nbv(f, 0, -1) + nbv(f, 1, 1) + nbv(f, 1, -1) + nbv(f, 2, 1) +
                nbv(f, 2, -1) - 6.0 * c[f]) * h2;
    };
    // curl of the 3-vector at base f0, component a (C:409-415)
    auto curl = [&](int f0, int a) -> double {
        if (a == 0)
            return (nbv(f0 + 2, 1, 1) - nbv(f0 + 2, 1, -1) - nbv(f0 + 1, 2, 1) + nbv(f0 + 1, 2, -1)) * h1;
        if (a == 1)
            return (nbv(f0 + 0, 2, 1) - nbv(f0 + 0, 2, -1) - nbv(f0 + 2, 0, 1) + nbv(f0 + 2, 0, -1)) * h1;
        return (nbv(f0 + 1, 0, 1) - nbv(f0 + 1, 0, -1) - nbv(f0 + 0, 1, 1) + nbv(f0 + 0, 1, -1)) * h1;
    };
#pragma unroll
    for (int a = 0; a < 3; a++) {
        double u = c[F_U + a], v = c[F_V + a];
        acc[F_U + a] = lap(F_U + a) - kp.MASS2 * u - 2.0 * Vp * u * rest[a] + kp.ETA * curl(F_TU, a);
        acc[F_V + a] = lap(F_V + a) - kp.MASS2 * v - 2.0 * Vp * v * rest[a] + kp.ETA * curl(F_TV, a);
        acc[F_TU + a] = lap(F_TU + a) - kp.MT2 * c[F_TU + a] + kp.ETA * curl(F_U, a);
        acc[F_TV + a] = lap(F_TV + a) - kp.MT2 * c[F_TV + a] + kp.ETA * curl(F_V, a);
    }
}

// Real-mode accelerations (compute_forces, C:743-922). AUX(f, idx) loads the
// pass-1 fields M[0..2] (f=0..2) and Q[0..2] (f=3..5) when alpha_cs/beta_h != 0.
template <class LD, class LA>
__device__ inline void forces_real(const KParams &kp, const double *c, const double *ip,
                                   const double *im, LD X, LA AUX, uint32_t ci, uint32_t ipx,
                                   uint32_t imx, uint32_t jp, uint32_t jm, uint32_t kp_,
                                   uint32_t km, double *acc) {
    const double h1 = kp.h1, h2 = kp.h2;
    const double MU = kp.MU, KAPPA = kp.KAPPA, MASS2 = kp.MASS2, ETA = kp.ETA;
    const double KH = kp.KH, ACS = kp.ACS, BH = kp.BH, KG = kp.KG;
    const int MODE = kp.rmode;
    auto nbv = [&](int f, int d, int sg) -> double {
        if (d == 0) return sg > 0 ? ip[f] : im[f];
        if (d == 1) return X(f, sg > 0 ? jp : jm);
        return X(f, sg > 0 ? kp_ : km);
    };
    auto nbi = [&](int d, int sg) -> uint32_t {
        if (d == 0) return sg > 0 ? ipx : imx;
        if (d == 1) return sg > 0 ? jp : jm;
        return sg > 0 ? kp_ : km;
    };
    auto curlF = [&](int f0, int a) -> double {
        if (a == 0)
            return (nbv(f0 + 2, 1, 1) - nbv(f0 + 2, 1, -1) - nbv(f0 + 1, 2, 1) + nbv(f0 + 1, 2, -1)) * h1;
        if (a == 1)
            return (nbv(f0 + 0, 2, 1) - nbv(f0 + 0, 2, -1) - nbv(f0 + 2, 0, 1) + nbv(f0 + 2, 0, -1)) * h1;
        return (nbv(f0 + 1, 0, 1) - nbv(f0 + 1, 0, -1) - nbv(f0 + 0, 1, 1) + nbv(f0 + 0, 1, -1)) * h1;
    };
    auto curlA = [&](int f0, int a) -> double {
        if (a == 0)
            return (AUX(f0 + 2, nbi(1, 1)) - AUX(f0 + 2, nbi(1, -1)) - AUX(f0 + 1, nbi(2, 1)) + AUX(f0 + 1, nbi(2, -1))) * h1;
        if (a == 1)
            return (AUX(f0 + 0, nbi(2, 1)) - AUX(f0 + 0, nbi(2, -1)) - AUX(f0 + 2, nbi(0, 1)) + AUX(f0 + 2, nbi(0, -1))) * h1;
        return (AUX(f0 + 1, nbi(0, 1)) - AUX(f0 + 1, nbi(0, -1)) - AUX(f0 + 0, nbi(1, 1)) + AUX(f0 + 0, nbi(1, -1))) * h1;
    };
    const int TU = 3;   // real mode: tu at fields 3..5
    double p0 = c[0], p1 = c[1], p2 = c[2];
    double sig = p0 * p0 + p1 * p1 + p2 * p2;
    double me2 = (MODE == 1) ? kp.IA / (1.0 + kp.IB * sig) : MASS2;
    double keff = (MODE == 3) ? KAPPA / (1.0 + KG * sig) : KAPPA;
    double P = p0 * p1 * p2, P2 = P * P;
    double den = 1.0 + keff * P2;
    double dVdP = MU * P / (den * den);
    double t2c = 0;
    if (MODE == 3 && KG > 0) {
        double D = 1.0 + KG * sig + KAPPA * P2;
        t2c = MU * KG * KAPPA * P2 * P2 / (D * D);
    }
    double cph[3] = {curlF(0, 0), curlF(0, 1), curlF(0, 2)};
    double hel = p0 * cph[0] + p1 * cph[1] + p2 * cph[2];
    double dP2[3] = {0, 0, 0};
    if (KH != 0) {
#pragma unroll
        for (int d = 0; d < 3; d++) {
            double Pp = nbv(0, d, 1) * nbv(1, d, 1) * nbv(2, d, 1);
            double Pm = nbv(0, d, -1) * nbv(1, d, -1) * nbv(2, d, -1);
            dP2[d] = (Pp * Pp - Pm * Pm) * h1;
        }
    }
    double cM[3] = {0, 0, 0}, cQ[3] = {0, 0, 0};
    if (ACS != 0) { cM[0] = curlA(0, 0); cM[1] = curlA(0, 1); cM[2] = curlA(0, 2); }
    if (BH != 0) { cQ[0] = curlA(3, 0); cQ[1] = curlA(3, 1); cQ[2] = curlA(3, 2); }
    double csq = cph[0] * cph[0] + cph[1] * cph[1] + cph[2] * cph[2];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        double lp = (nbv(a, 0, 1) + nbv(a, 0, -1) + nbv(a, 1, 1) + nbv(a, 1, -1) + nbv(a, 2, 1) +
                     nbv(a, 2, -1) - 6.0 * c[a]) * h2;
        double dPda = (a == 0) ? p1 * p2 : (a == 1) ? p0 * p2 : p0 * p1;
        double ct = curlF(TU, a);
        double chiral = 0;
        if (KH != 0) {
            double t1 = 2.0 * P * dPda * hel;
            double t2 = 2.0 * P2 * cph[a];
            double t3;
            if (a == 0) t3 = -(dP2[2] * p1 - dP2[1] * p2);
            else if (a == 1) t3 = -(dP2[0] * p2 - dP2[2] * p0);
            else t3 = -(dP2[1] * p0 - dP2[0] * p1);
            chiral = KH * (t1 + t2 + t3);
        }
        double cos_phi = -ACS * cM[a];
        double hard_phi = -2.0 * cQ[a];
        acc[a] = lp - me2 * c[a] - dVdP * dPda - t2c * c[a] + ETA * ct + chiral + cos_phi + hard_phi;
    }
#pragma unroll
    for (int a = 0; a < 3; a++) {
        double lt = (nbv(TU + a, 0, 1) + nbv(TU + a, 0, -1) + nbv(TU + a, 1, 1) + nbv(TU + a, 1, -1) +
                     nbv(TU + a, 2, 1) + nbv(TU + a, 2, -1) - 6.0 * c[TU + a]) * h2;
        double cp = curlF(0, a);
        double cos_th = 0;
        if (ACS != 0) cos_th = 2.0 * ACS * AUX(a, ci);
        double hard_th = -BH * csq * c[TU + a];
        acc[TU + a] = lt - kp.MT2 * c[TU + a] + ETA * cp + cos_th + hard_th;
    }
}

// Carried per-column gauge quantities from plane i-1.
struct GaugeCarry {
    double c0, s0;      // cos/sin th_0(x - e0)
    double p0, p2;      // plaquette sines (0,1) and (2,0) at x - e0
};

// Gauged accelerations (compute_forces_complex_gauge, C:1053-1263), one
// gauge, no Higgs. Writes acc[0..11] (matter, device order) and acc[12..14]
// (E_acc), and the carry for plane i+1.
template <class LD>
__device__ inline void forces_gauge(const KParams &kp, const double *c, const double *ip,
                                    const double *im, LD X, int N, int i, int j, int k,
                                    uint32_t jp, uint32_t jm, uint32_t kp_, uint32_t km,
                                    GaugeCarry &cy, double *acc) {
    const double h1 = kp.h1, h2 = kp.h2;
    auto nbv = [&](int f, int d, int sg) -> double {
        if (d == 0) return sg > 0 ? ip[f] : im[f];
        if (d == 1) return X(f, sg > 0 ? jp : jm);
        return X(f, sg > 0 ? kp_ : km);
    };
    // links and their cos/sin at x and x - e_d (C:1071-1079, 1144-1145)
    double cP[3], sP[3], cM[3], sM[3];
#pragma unroll
    for (int d = 0; d < 3; d++) sincos(c[F_TH + d], &sP[d], &cP[d]);
    cM[0] = cy.c0; sM[0] = cy.s0;
    sincos(nbv(F_TH + 1, 1, -1), &sM[1], &cM[1]);
    sincos(nbv(F_TH + 2, 2, -1), &sM[2], &cM[2]);

    // plaquettes (C:1088-1109) at x, x-e1, x-e2; x-e0 carried
    const int ip1 = wp(i, N), jm1 = wm(j, N), kp1 = wp(k, N), km1 = wm(k, N), jp1 = wp(j, N);
    double th0 = c[F_TH + 0], th1 = c[F_TH + 1], th2 = c[F_TH + 2];
    double px0 = plaq(th0, ip[F_TH + 1], nbv(F_TH + 0, 1, 1), th1);       // (0,1) at x
    double px1 = plaq(th1, nbv(F_TH + 2, 1, 1), nbv(F_TH + 1, 2, 1), th2); // (1,2) at x
    double px2 = plaq(th2, nbv(F_TH + 0, 2, 1), ip[F_TH + 2], th0);       // (2,0) at x
    // y = x - e1
    double t1_ym = nbv(F_TH + 1, 1, -1), t0_ym = nbv(F_TH + 0, 1, -1), t2_ym = nbv(F_TH + 2, 1, -1);
    double p0_ym = plaq(t0_ym, X(F_TH + 1, sidx(N, ip1, jm1, k)), th0, t1_ym);
    double p1_ym = plaq(t1_ym, th2, X(F_TH + 1, sidx(N, i, jm1, kp1)), t2_ym);
    // y = x - e2
    double t1_zm = nbv(F_TH + 1, 2, -1), t2_zm = nbv(F_TH + 2, 2, -1), t0_zm = nbv(F_TH + 0, 2, -1);
    double p1_zm = plaq(t1_zm, X(F_TH + 2, sidx(N, i, jp1, km1)), th1, t2_zm);
    double p2_zm = plaq(t2_zm, th0, X(F_TH + 2, sidx(N, ip1, j, km1)), t0_zm);
    double st[3];
    st[0] = (px0 - p0_ym) - (px2 - p2_zm);
    st[1] = -(px0 - cy.p0) + (px1 - p1_zm);
    st[2] = (px2 - cy.p2) - (px1 - p1_ym);

    // potential (C:1160-1166)
    double fu[6], fv[6];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        fu[a] = c[F_U + a]; fv[a] = c[F_V + a];
        fu[3 + a] = c[F_TU + a]; fv[3 + a] = c[F_TV + a];
    }
    double s2[3];
#pragma unroll
    for (int a = 0; a < 3; a++) s2[a] = fu[a] * fu[a] + fv[a] * fv[a];
    double s = s2[0] * s2[1] * s2[2];
    double den = 1.0 + kp.KAPPA * s;
    double Vp = 0.5 * kp.MU / (den * den);
    double rest[3] = {s2[1] * s2[2], s2[0] * s2[2], s2[0] * s2[1]};

    // transported neighbours, per direction (C:1134-1157), consumed as produced
    double lU[6], lV[6];          // Laplacian partial sums, per field f (Phi 0..2, Theta 0..2)
    double DU[3][6], DV[3][6];    // covariant central differences (only d != component used)
    const int C1[3] = {1, 2, 0}, C2[3] = {2, 0, 1};
#pragma unroll
    for (int d = 0; d < 3; d++) {
        double TRp[6], TIp[6], tup[3], tvp[3];
#pragma unroll
        for (int a = 0; a < 3; a++) {
#pragma unroll
            for (int h = 0; h < 2; h++) {       // h=0: Phi_a (u,v); h=1: Theta_a (tu,tv)
                const int fuix = h ? F_TU + a : F_U + a, fvix = h ? F_TV + a : F_V + a;
                double upn = nbv(fuix, d, 1), vpn = nbv(fvix, d, 1);
                double umn = nbv(fuix, d, -1), vmn = nbv(fvix, d, -1);
                double trp = cP[d] * upn - sP[d] * vpn, tip = cP[d] * vpn + sP[d] * upn;
                double trm = cM[d] * umn + sM[d] * vmn, tim = cM[d] * vmn - sM[d] * umn;
                const int f = h * 3 + a;
                if (d == 0) { lU[f] = trp + trm; lV[f] = tip + tim; }
                else { lU[f] = lU[f] + trp; lU[f] = lU[f] + trm; lV[f] = lV[f] + tip; lV[f] = lV[f] + tim; }
                DU[d][f] = (trp - trm) * h1;
                DV[d][f] = (tip - tim) * h1;
                TRp[f] = trp; TIp[f] = tip;
                if (h) { tup[a] = upn; tvp[a] = vpn; }
            }
        }
        // lattice current J_d and E-kick (C:1232-1250)
        double Jg = 0.0;
#pragma unroll
        for (int a = 0; a < 3; a++) {
            Jg += fu[a] * TIp[a] - fv[a] * TRp[a];
            Jg += fu[3 + a] * TIp[3 + a] - fv[3 + a] * TRp[3 + a];
        }
        const int d1 = C1[d], d2 = C2[d];
        double WpR1 = cP[d] * fu[d2] + sP[d] * fv[d2], WpI1 = cP[d] * fv[d2] - sP[d] * fu[d2];
        double WpR2 = cP[d] * fu[d1] + sP[d] * fv[d1], WpI2 = cP[d] * fv[d1] - sP[d] * fu[d1];
        double T12 = (fu[3 + d1] * TIp[d2] - fv[3 + d1] * TRp[d2]) + (tup[d1] * WpI1 - tvp[d1] * WpR1);
        double T21 = (fu[3 + d2] * TIp[d1] - fv[3 + d2] * TRp[d1]) + (tup[d2] * WpI2 - tvp[d2] * WpR2);
        double Jlat = kp.inva * Jg - 0.5 * kp.ETA * (T12 - T21);
        acc[F_TH + d] = kp.STP * st[d] + kp.GG * Jlat;
    }
    // matter accelerations (C:1179-1204)
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const int d1 = C1[a], d2 = C2[a];
        double reDxT = DU[d1][3 + C2[a]] - DU[d2][3 + C1[a]];
        double imDxT = DV[d1][3 + C2[a]] - DV[d2][3 + C1[a]];
        double reDxP = DU[d1][C2[a]] - DU[d2][C1[a]];
        double imDxP = DV[d1][C2[a]] - DV[d2][C1[a]];
        double lUP = (lU[a] - 6.0 * fu[a]) * h2;
        double lVP = (lV[a] - 6.0 * fv[a]) * h2;
        double lUT = (lU[3 + a] - 6.0 * fu[3 + a]) * h2;
        double lVT = (lV[3 + a] - 6.0 * fv[3 + a]) * h2;
        acc[F_U + a] = lUP - kp.MASS2 * fu[a] - 2.0 * Vp * fu[a] * rest[a] + kp.ETA * reDxT;
        acc[F_V + a] = lVP - kp.MASS2 * fv[a] - 2.0 * Vp * fv[a] * rest[a] + kp.ETA * imDxT;
        acc[F_TU + a] = lUT - kp.MT2 * fu[3 + a] + kp.ETA * reDxP;
        acc[F_TV + a] = lVT - kp.MT2 * fv[3 + a] + kp.ETA * imDxP;
    }
    cy.c0 = cP[0]; cy.s0 = sP[0]; cy.p0 = px0; cy.p2 = px2;
}

// carry for the segment's first plane: quantities at (i-1, j, k)
template <class LD>
__device__ inline void gauge_prologue(const KParams &kp, const double *cm, const double *c0, LD X,
                                      int N, int im1, int j, int k, GaugeCarry &cy) {
    sincos(cm[F_TH + 0], &cy.s0, &cy.c0);
    const int jp1 = wp(j, N), kp1 = wp(k, N);
    cy.p0 = plaq(cm[F_TH + 0], c0[F_TH + 1], X(F_TH + 0, sidx(N, im1, jp1, k)), cm[F_TH + 1]);
    cy.p2 = plaq(cm[F_TH + 2], X(F_TH + 0, sidx(N, im1, j, kp1)), c0[F_TH + 2], cm[F_TH + 0]);
    (void)kp;
}

// sponge factor of scp_sim.c:1447-1453 (1.0 and false outside the sponge)
__device__ inline bool sponge(const KParams &kp, int i, int j, int k, double &d) {
    double x = -kp.L + i * kp.dx, y = -kp.L + j * kp.dx, z = -kp.L + k * kp.dx;
    double r = sqrt(x * x + y * y + z * z);
    if (r > kp.Rd) {
        double s = (r - kp.Rd) / kp.DW;
        if (s > 1) s = 1;
        d = 1.0 - kp.DR * s * s;
        return true;
    }
    return false;
}

// ---------------------------------------------------------------------------
// the fused step kernel, FP64 storage
// ---------------------------------------------------------------------------

template <int MODE, int KIND, bool EMIT>
__global__ void __launch_bounds__(TK * TJ)
k_step64(KParams kp, Bufs<double> b, int DI) {
    constexpr int NP = MODE == 2 ? 15 : (MODE == 1 ? 12 : 6);
    const int N = kp.N;
    const Geo g = geometry(N, DI);
    if (!g.valid || g.i0 >= N) return;
    const uint32_t fs = kp.fs;
    const double *__restrict__ src = (KIND == K_STEP) ? b.pin : b.xo;
    auto X = [&](int f, uint32_t idx) -> double { return src[(size_t)f * fs + idx]; };
    auto AUX = [&](int f, uint32_t idx) -> double { return b.aux[(size_t)f * fs + idx]; };
    const int j = g.j, k = g.k;
    const int jp1 = wp(j, N), jm1 = wm(j, N), kp1 = wp(k, N), km1 = wm(k, N);

    double cm[NP], c0[NP], cp[NP];
    {
        const int im1 = wm(g.i0, N);
        const uint32_t a = sidx(N, im1, j, k), c = sidx(N, g.i0, j, k);
#pragma unroll
        for (int f = 0; f < NP; f++) { cm[f] = X(f, a); c0[f] = X(f, c); }
    }
    GaugeCarry cy{};
    if constexpr (MODE == 2) gauge_prologue(kp, cm, c0, X, N, wm(g.i0, N), j, k, cy);

    for (int i = g.i0; i < g.i1; i++) {
        const int ip1 = wp(i, N), im1 = wm(i, N);
        const uint32_t ci = sidx(N, i, j, k);
        {
            const uint32_t a = sidx(N, ip1, j, k);
#pragma unroll
            for (int f = 0; f < NP; f++) cp[f] = X(f, a);
        }
        const uint32_t jp = sidx(N, i, jp1, k), jm = sidx(N, i, jm1, k);
        const uint32_t kpx = sidx(N, i, j, kp1), kmx = sidx(N, i, j, km1);
        double acc[NP];
        if constexpr (MODE == 2)
            forces_gauge(kp, c0, cp, cm, X, N, i, j, k, jp, jm, kpx, kmx, cy, acc);
        else if constexpr (MODE == 1)
            forces_complex(kp, c0, cp, cm, X, jp, jm, kpx, kmx, acc);
        else
            forces_real(kp, c0, cp, cm, X, AUX, ci, sidx(N, ip1, j, k), sidx(N, im1, j, k), jp, jm,
                        kpx, kmx, acc);

        if constexpr (KIND == K_FORCES) {
#pragma unroll
            for (int f = 0; f < NP; f++) b.acc[(size_t)f * fs + ci] = acc[f];
        } else {
            double dmp = 1.0;
            const bool damp = (KIND == K_STEP) && kp.damp && sponge(kp, i, j, k, dmp);
#pragma unroll
            for (int f = 0; f < NP; f++) {
                const size_t o = (size_t)f * fs + ci;
                double v;
                if constexpr (KIND == K_INIT) {
                    v = b.vo[o];
                } else {
                    v = __builtin_nontemporal_load(&b.w[o]);
                    v += kp.hdt * acc[f];                  // stage 4 (C:1680-1696)
                    if (damp) v *= dmp;                    // stage 5 (C:1454-1460)
                }
                double w = v + kp.hdt * acc[f];            // next stage 1 (C:1597-1613)
                double x = c0[f];
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
                    __builtin_nontemporal_store(x, &b.xo[o]);
                    __builtin_nontemporal_store(v, &b.vo[o]);
                }
            }
        }
#pragma unroll
        for (int f = 0; f < NP; f++) { cm[f] = c0[f]; c0[f] = cp[f]; }
    }
}

// Real-mode pass 1 (C:755-786): Cosserat mismatch and hardening vector.
template <typename T>
__global__ void k_real_aux(KParams kp, const T *pin, double *aux) {
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

int launch_real_aux(const KParams &kp, const double *pin, double *aux, hipStream_t s) {
    hipLaunchKernelGGL((k_real_aux<double>), dim3((kp.n3 + 255) / 256), dim3(256), 0, s, kp, pin, aux);
    return (int)hipGetLastError();
}

}  // namespace scp
