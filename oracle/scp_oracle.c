/* scp_oracle.c — CPU FP64 restatement of the reference sfa/sim step (the oracle).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for tests/, smoke() and bench.py's
 * cpu_baseline leg. The shipped product never links or calls this file.
 *
 * Every function restates one reference function of sfa/sim/scp_sim.c and
 * keeps its floating-point expression trees operation for operation, so that
 * with FMA contraction off (see Makefile) each per-site result rounds exactly
 * as the reference does. Reductions (diagnostics) are summed sequentially in
 * site order; the reference uses OpenMP reductions, so those agree to the last
 * few ulps, not bit for bit.
 *
 * Parity pins (tests/test_oracle.py):
 *   or_forces/or_step  — bitwise vs the reference binary built from source
 *                        (oracle/_ref, fixtures in tests/golden/), real-limit
 *                        identity, uniform-field ODE, Q conservation at eta=0.
 *   or_diag            — reference diag.tsv rows to printed precision; closed
 *                        forms (Gaussian mass integral, plane-wave gradient).
 *   or_gauss_project   — post-condition max|G-Gbar| <= 1e-12 and the
 *                        reference's frame-0 E field.
 *   or_gauge_transform — gauge invariance of |acc|^2, E_acc, G (<= 1e-12).
 *   or_qflow           — the fixed-charge flow of sfa/seed/eta_qflow.c: its f32
 *                        output and every printed row, bitwise, against that
 *                        program built from source (oracle/_ref/eta_qflow,
 *                        one thread, so its reductions run in site order).
 */
#include "scp_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define OR_PI 3.14159265358979323846   /* scp_config.h:16 */

int or_ncols(int mode) { return mode == 2 ? 30 : (mode == 1 ? 24 : 12); }
int or_nacc(int mode)  { return mode == 2 ? 15 : (mode == 1 ? 12 : 6); }

/* Column views into the SFA-ordered state buffer. */
typedef struct {
    const double *u[3], *tu[3], *ud[3], *tud[3];
    const double *v[3], *tv[3], *vd[3], *tvd[3];
    const double *th[3], *E[3];
} View;

static View make_view(const double *s, long n3, int mode) {
    View w; memset(&w, 0, sizeof w);
    for (int a = 0; a < 3; a++) {
        w.u[a] = s + (0 + a) * n3;   w.tu[a] = s + (3 + a) * n3;
        w.ud[a] = s + (6 + a) * n3;  w.tud[a] = s + (9 + a) * n3;
        if (mode >= 1) {
            w.v[a] = s + (12 + a) * n3;  w.tv[a] = s + (15 + a) * n3;
            w.vd[a] = s + (18 + a) * n3; w.tvd[a] = s + (21 + a) * n3;
        }
        if (mode == 2) { w.th[a] = s + (24 + a) * n3; w.E[a] = s + (27 + a) * n3; }
    }
    return w;
}

/* Periodic neighbour indices of site idx (scp_sim.c:941-946). nbr[d][0] = +d, nbr[d][1] = -d. */
static inline void neighbours(int N, long idx, long np[3], long nm[3]) {
    const long NN = (long)N * N;
    int i = (int)(idx / NN), j = (int)((idx / N) % N), k = (int)(idx % N);
    np[0] = (long)((i + 1) % N) * NN + (long)j * N + k;
    nm[0] = (long)((i - 1 + N) % N) * NN + (long)j * N + k;
    np[1] = (long)i * NN + (long)((j + 1) % N) * N + k;
    nm[1] = (long)i * NN + (long)((j - 1 + N) % N) * N + k;
    np[2] = (long)i * NN + (long)j * N + (k + 1) % N;
    nm[2] = (long)i * NN + (long)j * N + (k - 1 + N) % N;
}

/* curl of a 3-vector field at a site, central differences (scp_sim.c:409-415) */
static inline double curl_at(const double *const F[3], int a, const long np[3], const long nm[3], double h1) {
    if (a == 0) return (F[2][np[1]] - F[2][nm[1]] - F[1][np[2]] + F[1][nm[2]]) * h1;
    if (a == 1) return (F[0][np[2]] - F[0][nm[2]] - F[2][np[0]] + F[2][nm[0]]) * h1;
    return (F[1][np[0]] - F[1][nm[0]] - F[0][np[1]] + F[0][nm[1]]) * h1;
}

static inline double lap_at(const double *F, long c, const long np[3], const long nm[3], double h2) {
    return (F[np[0]] + F[nm[0]] + F[np[1]] + F[nm[1]] + F[np[2]] + F[nm[2]] - 6.0 * F[c]) * h2;
}

/* ---------------- forces ---------------- */

/* Real mode, scp_sim.c:743-922 (compute_forces), including the CPU-implemented
 * optional terms: mass modes 1/3, chiral kappa_h, Cosserat alpha_cs (two-pass),
 * hardening beta_h (two-pass). */
static void gpu_terms(const OrParams *p, const double *state, double *acc);

static void forces_real(const OrParams *p, const double *state, double *acc) {
    const int N = p->N; const long n3 = (long)N * N * N;
    const double h2 = 1.0 / (p->dx * p->dx), h1 = 1.0 / (2.0 * p->dx);
    const double MU = p->mu, KAPPA = p->kappa, MASS2 = p->m2, MT2 = p->mtheta2, ETA = p->eta;
    const double KH = p->kappa_h, ACS = p->alpha_cs, BH = p->beta_h;
    const int MODE = p->rmode;
    const double KG = p->kappa_gamma, IA = p->inv_alpha, IB = p->inv_beta;
    View w = make_view(state, n3, 0);
    double *M[3] = {0}, *Qh[3] = {0};
    for (int a = 0; a < 3; a++) { M[a] = calloc(n3, sizeof(double)); Qh[a] = calloc(n3, sizeof(double)); }

    if (ACS != 0 || BH != 0) {                              /* pass 1: C:755-786 */
        #pragma omp parallel for schedule(static)
        for (long c = 0; c < n3; c++) {
            long np[3], nm[3]; neighbours(N, c, np, nm);
            double cp[3];
            for (int a = 0; a < 3; a++) cp[a] = curl_at(w.u, a, np, nm, h1);
            if (ACS != 0) for (int a = 0; a < 3; a++) M[a][c] = cp[a] * 0.5 - w.tu[a][c];
            if (BH != 0) {
                double t0 = w.tu[0][c], t1 = w.tu[1][c], t2 = w.tu[2][c];
                double T2 = t0 * t0 + t1 * t1 + t2 * t2;
                double coeff = 0.5 * BH * T2;
                for (int a = 0; a < 3; a++) Qh[a][c] = coeff * cp[a];
            }
        }
    }
    const double *Mc[3] = {M[0], M[1], M[2]}, *Qc[3] = {Qh[0], Qh[1], Qh[2]};

    #pragma omp parallel for schedule(static)
    for (long c = 0; c < n3; c++) {                          /* pass 2: C:789-921 */
        long np[3], nm[3]; neighbours(N, c, np, nm);
        double p0 = w.u[0][c], p1 = w.u[1][c], p2 = w.u[2][c];
        double sig = p0 * p0 + p1 * p1 + p2 * p2;
        double me2 = (MODE == 1) ? IA / (1.0 + IB * sig) : MASS2;
        double keff = (MODE == 3) ? KAPPA / (1.0 + KG * sig) : KAPPA;
        double P = p0 * p1 * p2, P2 = P * P;
        double den = 1.0 + keff * P2;
        double dVdP = MU * P / (den * den);
        double t2c = 0;
        if (MODE == 3 && KG > 0) {
            double D = 1.0 + KG * sig + KAPPA * P2;
            t2c = MU * KG * KAPPA * P2 * P2 / (D * D);
        }
        double cph[3];
        for (int a = 0; a < 3; a++) cph[a] = curl_at(w.u, a, np, nm, h1);
        double hel = p0 * cph[0] + p1 * cph[1] + p2 * cph[2];
        double dP2[3] = {0, 0, 0};
        if (KH != 0) {
            double Pn[3][2];
            for (int d = 0; d < 3; d++) {
                Pn[d][0] = w.u[0][np[d]] * w.u[1][np[d]] * w.u[2][np[d]];
                Pn[d][1] = w.u[0][nm[d]] * w.u[1][nm[d]] * w.u[2][nm[d]];
            }
            for (int d = 0; d < 3; d++) dP2[d] = (Pn[d][0] * Pn[d][0] - Pn[d][1] * Pn[d][1]) * h1;
        }
        double cM[3] = {0, 0, 0}, cQ[3] = {0, 0, 0};
        if (ACS != 0) for (int a = 0; a < 3; a++) cM[a] = curl_at(Mc, a, np, nm, h1);
        if (BH != 0)  for (int a = 0; a < 3; a++) cQ[a] = curl_at(Qc, a, np, nm, h1);
        double csq = cph[0] * cph[0] + cph[1] * cph[1] + cph[2] * cph[2];
        for (int a = 0; a < 3; a++) {
            double lap = lap_at(w.u[a], c, np, nm, h2);
            double dPda = (a == 0) ? p1 * p2 : (a == 1) ? p0 * p2 : p0 * p1;
            double ct = curl_at(w.tu, a, np, nm, h1);
            double chiral = 0;
            if (KH != 0) {
                double t1 = 2.0 * P * dPda * hel;
                double t2 = 2.0 * P2 * cph[a];
                double t3;
                if (a == 0)      t3 = -(dP2[2] * p1 - dP2[1] * p2);
                else if (a == 1) t3 = -(dP2[0] * p2 - dP2[2] * p0);
                else             t3 = -(dP2[1] * p0 - dP2[0] * p1);
                chiral = KH * (t1 + t2 + t3);
            }
            double cos_phi = -ACS * cM[a];
            double hard_phi = -2.0 * cQ[a];
            acc[(0 + a) * n3 + c] = lap - me2 * w.u[a][c] - dVdP * dPda
                                  - t2c * w.u[a][c] + ETA * ct + chiral
                                  + cos_phi + hard_phi;
        }
        for (int a = 0; a < 3; a++) {
            double lapt = lap_at(w.tu[a], c, np, nm, h2);
            double cp = curl_at(w.u, a, np, nm, h1);
            double cos_th = 0;
            if (ACS != 0) cos_th = 2.0 * ACS * M[a][c];
            double hard_th = -BH * csq * w.tu[a][c];
            acc[(3 + a) * n3 + c] = lapt - MT2 * w.tu[a][c] + ETA * cp + cos_th + hard_th;
        }
    }
    for (int a = 0; a < 3; a++) { free(M[a]); free(Qh[a]); }
    gpu_terms(p, state, acc);
}

/* The real-mode terms the reference applies in its GPU force kernel only
 * (compute_forces_kernel, scp_sim.cu:1144-1248), restated per site in that
 * kernel's order, added onto the accelerations of compute_forces. The theta
 * velocities are the state's (the half-kicked ones inside a step, as there).
 * The frequency term uses omega = |theta_vel| / (|theta| + 0.01) at the site
 * and its six neighbours (ip, im, jp, jm, kp, km). */
static double omega_at(const View *w, long c) {
    double tv = w->tud[0][c] * w->tud[0][c] + w->tud[1][c] * w->tud[1][c] + w->tud[2][c] * w->tud[2][c];
    double th = w->tu[0][c] * w->tu[0][c] + w->tu[1][c] * w->tu[1][c] + w->tu[2][c] * w->tu[2][c];
    return sqrt(tv) / (sqrt(th) + 0.01);
}
static void gpu_terms(const OrParams *p, const double *state, double *acc) {
    const int on = (p->theta_sat > 0 && p->gamma_conv > 0) || p->sigma_grad > 0 || p->sigma_cubic > 0 ||
                   p->sigma_freq > 0 || p->sigma_cross > 0 || p->lambda_self > 0;
    if (!on) return;
    const int N = p->N; const long n3 = (long)N * N * N;
    const double h1 = 1.0 / (2.0 * p->dx);
    View w = make_view(state, n3, 0);
    #pragma omp parallel for schedule(static)
    for (long c = 0; c < n3; c++) {
        long np[3], nm[3]; neighbours(N, c, np, nm);
        double *ap[3], *at[3];
        for (int a = 0; a < 3; a++) { ap[a] = acc + (0 + a) * n3 + c; at[a] = acc + (3 + a) * n3 + c; }
        const double t0 = w.tu[0][c], t1 = w.tu[1][c], t2 = w.tu[2][c];
        const double th[3] = {t0, t1, t2};
        double ct[3];
        for (int a = 0; a < 3; a++) ct[a] = curl_at(w.tu, a, np, nm, h1);
        if (p->theta_sat > 0 && p->gamma_conv > 0) {
            double th2 = t0 * t0 + t1 * t1 + t2 * t2;
            double excess = th2 - p->theta_sat * p->theta_sat;
            if (excess > 0)
                for (int a = 0; a < 3; a++) {
                    *at[a] -= p->gamma_conv * excess * th[a];
                    *ap[a] += p->gamma_conv * excess * ct[a];
                }
        }
        if (p->sigma_grad > 0) {
            double g2 = 0;
            for (int a = 0; a < 3; a++) {
                double dx = (w.tu[a][np[0]] - w.tu[a][nm[0]]) * h1;
                double dy = (w.tu[a][np[1]] - w.tu[a][nm[1]]) * h1;
                double dz = (w.tu[a][np[2]] - w.tu[a][nm[2]]) * h1;
                g2 += dx * dx + dy * dy + dz * dz;
            }
            double th2 = t0 * t0 + t1 * t1 + t2 * t2;
            double h = t0 * ct[0] + t1 * ct[1] + t2 * ct[2];
            double se = p->sigma_grad * (1.0 + p->chi_chiral * h / (th2 + 0.01));
            if (se < 0) se = 0;
            for (int a = 0; a < 3; a++) {
                *at[a] -= se * g2 * th[a];
                *ap[a] += se * g2 * ct[a];
            }
        }
        if (p->sigma_cubic > 0) {
            double th2 = t0 * t0 + t1 * t1 + t2 * t2;
            for (int a = 0; a < 3; a++) {
                *at[a] -= p->sigma_cubic * th2 * th[a];
                *ap[a] += p->sigma_cubic * th2 * ct[a];
            }
        }
        if (p->sigma_freq > 0) {
            double oc = omega_at(&w, c), dw = 0;
            long nb[6] = {np[0], nm[0], np[1], nm[1], np[2], nm[2]};
            for (int q = 0; q < 6; q++) dw += fabs(oc - omega_at(&w, nb[q]));
            double dwa = dw / 6;
            for (int a = 0; a < 3; a++) {
                *at[a] -= p->sigma_freq * dwa * th[a];
                *ap[a] += p->sigma_freq * dwa * ct[a];
            }
        }
        if (p->sigma_cross > 0) {
            double th2 = t0 * t0 + t1 * t1 + t2 * t2;
            double ph2 = w.u[0][c] * w.u[0][c] + w.u[1][c] * w.u[1][c] + w.u[2][c] * w.u[2][c];
            for (int a = 0; a < 3; a++) {
                *ap[a] -= p->sigma_cross * th2 * w.u[a][c];
                *at[a] -= p->sigma_cross * ph2 * th[a];
            }
        }
        if (p->lambda_self > 0) {
            double th2 = t0 * t0 + t1 * t1 + t2 * t2;
            for (int a = 0; a < 3; a++) *at[a] -= p->lambda_self * th2 * th[a];
        }
    }
}

/* Complex mode, scp_sim.c:931-994 (compute_forces_complex). */
static void forces_complex(const OrParams *p, const double *state, double *acc) {
    const int N = p->N; const long n3 = (long)N * N * N;
    const double h2 = 1.0 / (p->dx * p->dx), h1 = 1.0 / (2.0 * p->dx);
    const double MU = p->mu, KAPPA = p->kappa, MASS2 = p->m2, MT2 = p->mtheta2, ETA = p->eta;
    View w = make_view(state, n3, 1);
    #pragma omp parallel for schedule(static)
    for (long c = 0; c < n3; c++) {
        long np[3], nm[3]; neighbours(N, c, np, nm);
        double s2[3];
        for (int a = 0; a < 3; a++) s2[a] = w.u[a][c] * w.u[a][c] + w.v[a][c] * w.v[a][c];
        double s = s2[0] * s2[1] * s2[2];
        double den = 1.0 + KAPPA * s;
        double Vp = 0.5 * MU / (den * den);
        double rest[3] = { s2[1] * s2[2], s2[0] * s2[2], s2[0] * s2[1] };   /* never s/s2_a (C:957) */
        for (int a = 0; a < 3; a++) {
            double lu = lap_at(w.u[a], c, np, nm, h2);
            double ctu = curl_at(w.tu, a, np, nm, h1);
            acc[(0 + a) * n3 + c] = lu - MASS2 * w.u[a][c] - 2.0 * Vp * w.u[a][c] * rest[a] + ETA * ctu;
            double lv = lap_at(w.v[a], c, np, nm, h2);
            double ctv = curl_at(w.tv, a, np, nm, h1);
            acc[(6 + a) * n3 + c] = lv - MASS2 * w.v[a][c] - 2.0 * Vp * w.v[a][c] * rest[a] + ETA * ctv;
            double ltu = lap_at(w.tu[a], c, np, nm, h2);
            double cu = curl_at(w.u, a, np, nm, h1);
            acc[(3 + a) * n3 + c] = ltu - MT2 * w.tu[a][c] + ETA * cu;
            double ltv = lap_at(w.tv[a], c, np, nm, h2);
            double cv = curl_at(w.v, a, np, nm, h1);
            acc[(9 + a) * n3 + c] = ltv - MT2 * w.tv[a][c] + ETA * cv;
        }
    }
}

/* sin of the plaquette angle in plane p at site y; planes (0,1),(1,2),(2,0)
 * (scp_sim.c:1088-1109): th_a(y) + th_b(y+a) - th_a(y+b) - th_b(y). */
static inline double plaq_sin(const View *w, int N, long y, int pl) {
    static const int PA[3] = {0, 1, 2}, PB[3] = {1, 2, 0};
    int a = PA[pl], b = PB[pl];
    long np[3], nm[3]; neighbours(N, y, np, nm);
    double ang = w->th[a][y] + w->th[b][np[a]] - w->th[a][np[b]] - w->th[b][y];
    return sin(ang);
}

/* Gauged mode, scp_sim.c:1053-1263 (compute_forces_complex_gauge), single
 * gauge, no Higgs sector. */
static void forces_gauge(const OrParams *p, const double *state, double *acc) {
    const int N = p->N; const long n3 = (long)N * N * N;
    const double dx = p->dx, h2 = 1.0 / (dx * dx), h1 = 1.0 / (2.0 * dx), inva = 1.0 / dx;
    const double MU = p->mu, KAPPA = p->kappa, MASS2 = p->m2, MT2 = p->mtheta2, ETA = p->eta;
    const double GG = p->g, STP = 1.0 / (GG * dx * dx * dx);
    View w = make_view(state, n3, 2);
    double *lc[3], *ls[3], *ps[3];
    for (int d = 0; d < 3; d++) {
        lc[d] = malloc(n3 * sizeof(double)); ls[d] = malloc(n3 * sizeof(double));
        ps[d] = malloc(n3 * sizeof(double));
    }
    #pragma omp parallel for schedule(static)
    for (long c = 0; c < n3; c++)                              /* pass A: C:1071-1079 */
        for (int d = 0; d < 3; d++) { lc[d][c] = cos(w.th[d][c]); ls[d][c] = sin(w.th[d][c]); }
    #pragma omp parallel for schedule(static)
    for (long c = 0; c < n3; c++)                              /* pass B: C:1088-1109 */
        for (int pl = 0; pl < 3; pl++) ps[pl][c] = plaq_sin(&w, N, c, pl);

    static const int C1[3] = {1, 2, 0}, C2[3] = {2, 0, 1};
    #pragma omp parallel for schedule(static)
    for (long c = 0; c < n3; c++) {                            /* pass C: C:1111-1262 */
        long np[3], nm[3]; neighbours(N, c, np, nm);
        double fu[6], fv[6];
        for (int a = 0; a < 3; a++) {
            fu[a] = w.u[a][c];  fv[a] = w.v[a][c];
            fu[3 + a] = w.tu[a][c]; fv[3 + a] = w.tv[a][c];
        }
        double TRp[3][6], TIp[3][6], TRm[3][6], TIm[3][6], tup[3][3], tvp[3][3];
        for (int d = 0; d < 3; d++) {
            double cP = lc[d][c], sP = ls[d][c], cM = lc[d][nm[d]], sM = ls[d][nm[d]];
            for (int a = 0; a < 3; a++) {
                double upn = w.u[a][np[d]], vpn = w.v[a][np[d]];
                double umn = w.u[a][nm[d]], vmn = w.v[a][nm[d]];
                double tpn = w.tu[a][np[d]], wpn = w.tv[a][np[d]];
                double tmn = w.tu[a][nm[d]], wmn = w.tv[a][nm[d]];
                TRp[d][a] = cP * upn - sP * vpn;      TIp[d][a] = cP * vpn + sP * upn;
                TRm[d][a] = cM * umn + sM * vmn;      TIm[d][a] = cM * vmn - sM * umn;
                TRp[d][3 + a] = cP * tpn - sP * wpn;  TIp[d][3 + a] = cP * wpn + sP * tpn;
                TRm[d][3 + a] = cM * tmn + sM * wmn;  TIm[d][3 + a] = cM * wmn - sM * tmn;
                tup[d][a] = tpn; tvp[d][a] = wpn;
            }
        }
        double s2[3];
        for (int a = 0; a < 3; a++) s2[a] = fu[a] * fu[a] + fv[a] * fv[a];
        double s = s2[0] * s2[1] * s2[2];
        double den = 1.0 + KAPPA * s;
        double Vp = 0.5 * MU / (den * den);
        double rest[3] = { s2[1] * s2[2], s2[0] * s2[2], s2[0] * s2[1] };
        double DU[3][6], DV[3][6];
        for (int d = 0; d < 3; d++)
            for (int f = 0; f < 6; f++) {
                DU[d][f] = (TRp[d][f] - TRm[d][f]) * h1;
                DV[d][f] = (TIp[d][f] - TIm[d][f]) * h1;
            }
        for (int a = 0; a < 3; a++) {
            int d1 = C1[a], d2 = C2[a];
            double reDxT = DU[d1][3 + C2[a]] - DU[d2][3 + C1[a]];
            double imDxT = DV[d1][3 + C2[a]] - DV[d2][3 + C1[a]];
            double reDxP = DU[d1][C2[a]] - DU[d2][C1[a]];
            double imDxP = DV[d1][C2[a]] - DV[d2][C1[a]];
            double lUP = (TRp[0][a] + TRm[0][a] + TRp[1][a] + TRm[1][a] + TRp[2][a] + TRm[2][a] - 6.0 * fu[a]) * h2;
            double lVP = (TIp[0][a] + TIm[0][a] + TIp[1][a] + TIm[1][a] + TIp[2][a] + TIm[2][a] - 6.0 * fv[a]) * h2;
            double lUT = (TRp[0][3 + a] + TRm[0][3 + a] + TRp[1][3 + a] + TRm[1][3 + a] + TRp[2][3 + a] + TRm[2][3 + a] - 6.0 * fu[3 + a]) * h2;
            double lVT = (TIp[0][3 + a] + TIm[0][3 + a] + TIp[1][3 + a] + TIm[1][3 + a] + TIp[2][3 + a] + TIm[2][3 + a] - 6.0 * fv[3 + a]) * h2;
            acc[(0 + a) * n3 + c] = lUP - MASS2 * fu[a] - 2.0 * Vp * fu[a] * rest[a] + ETA * reDxT;
            acc[(6 + a) * n3 + c] = lVP - MASS2 * fv[a] - 2.0 * Vp * fv[a] * rest[a] + ETA * imDxT;
            acc[(3 + a) * n3 + c] = lUT - MT2 * fu[3 + a] + ETA * reDxP;
            acc[(9 + a) * n3 + c] = lVT - MT2 * fv[3 + a] + ETA * imDxP;
        }
        double st[3];
        st[0] =  (ps[0][c] - ps[0][nm[1]]) - (ps[2][c] - ps[2][nm[2]]);
        st[1] = -(ps[0][c] - ps[0][nm[0]]) + (ps[1][c] - ps[1][nm[2]]);
        st[2] =  (ps[2][c] - ps[2][nm[0]]) - (ps[1][c] - ps[1][nm[1]]);
        for (int d = 0; d < 3; d++) {
            double cP = lc[d][c], sP = ls[d][c];
            double Jg = 0.0;
            for (int a = 0; a < 3; a++) {
                Jg += fu[a] * TIp[d][a] - fv[a] * TRp[d][a];
                Jg += fu[3 + a] * TIp[d][3 + a] - fv[3 + a] * TRp[d][3 + a];
            }
            int d1 = C1[d], d2 = C2[d];
            double WpR1 = cP * fu[d2] + sP * fv[d2], WpI1 = cP * fv[d2] - sP * fu[d2];
            double WpR2 = cP * fu[d1] + sP * fv[d1], WpI2 = cP * fv[d1] - sP * fu[d1];
            double T12 = (fu[3 + d1] * TIp[d][d2] - fv[3 + d1] * TRp[d][d2])
                       + (tup[d][d1] * WpI1 - tvp[d][d1] * WpR1);
            double T21 = (fu[3 + d2] * TIp[d][d1] - fv[3 + d2] * TRp[d][d1])
                       + (tup[d][d2] * WpI2 - tvp[d][d2] * WpR2);
            double Jlat = inva * Jg - 0.5 * ETA * (T12 - T21);
            acc[(12 + d) * n3 + c] = STP * st[d] + GG * Jlat;
        }
    }
    for (int d = 0; d < 3; d++) { free(lc[d]); free(ls[d]); free(ps[d]); }
}

void or_forces(const OrParams *p, const double *state, double *acc) {
    if (p->mode == 2 && p->g != 0.0) forces_gauge(p, state, acc);
    else if (p->mode >= 1) forces_complex(p, state, acc);     /* g == 0 dispatch, C:1669-1670 */
    else forces_real(p, state, acc);
}

/* ---------------- integrator ---------------- */

/* Spherical sponge on velocities and E (scp_sim.c:1440-1473). */
static void damp(const OrParams *p, double *state) {
    const int N = p->N; const long n3 = (long)N * N * N, NN = (long)N * N;
    const double L = p->L, dx = p->dx, DW = p->damp_width, DR = p->damp_rate;
    if (DW <= 0 || DR <= 0) return;
    const double R = L - DW;
    const int nc = or_ncols(p->mode);
    #pragma omp parallel for schedule(static)
    for (long c = 0; c < n3; c++) {
        int i = (int)(c / NN), j = (int)((c / N) % N), k = (int)(c % N);
        double x = -L + i * dx, y = -L + j * dx, z = -L + k * dx;
        double r = sqrt(x * x + y * y + z * z);
        if (r > R) {
            double s = (r - R) / DW; if (s > 1) s = 1;
            double f = 1.0 - DR * s * s;
            for (int col = 0; col < nc; col++) {
                int vel = (col >= 6 && col < 12) || (col >= 18 && col < 24) || col >= 27;
                if (vel) state[col * n3 + c] *= f;
            }
        }
    }
}

/* Velocity column of each acceleration block, in state-column numbering. */
static int acc_to_vel(int b) {
    static const int T[15] = {6, 7, 8, 9, 10, 11, 18, 19, 20, 21, 22, 23, 27, 28, 29};
    return T[b];
}
static int vel_to_pos(int vc) {
    if (vc < 12) return vc - 6;
    if (vc < 24) return vc - 6;
    return vc - 3;       /* E_d (27+d) drives th_d (24+d) */
}

/* apply_gradient_bc, scp_sim.c:1480-1555 (stage 5 of bc 1, real and complex
 * modes): the x-slabs i < M and i >= N-M return to the pinned positions and
 * velocities with zero acceleration; then, for i in [M, N-M), the rows j = 0
 * and j = N-1 of every matter position and velocity column are extrapolated
 * linearly from the two rows inside. */
static void gradient_bc(const OrParams *p, double *state, double *acc) {
    const int N = p->N, M = p->gradient_margin;
    const long NN = (long)N * N, n3 = NN * N;
    const int ncol = p->mode >= 1 ? 24 : 12, nacc = p->mode >= 1 ? 12 : 6;
    for (long c = 0; c < n3; c++) {
        const int i = (int)(c / NN);
        if (i < M || i >= N - M) {
            for (int col = 0; col < ncol; col++) state[col * n3 + c] = p->pin[col * n3 + c];
            for (int b = 0; b < nacc; b++) acc[b * n3 + c] = 0;
        }
    }
    for (int i = M; i < N - M; i++)
        for (int k = 0; k < N; k++)
            for (int col = 0; col < ncol; col++) {
                double *f = state + col * n3 + (long)i * NN + k;
                f[0] = 2 * f[N] - f[2L * N];
                f[(long)(N - 1) * N] = 2 * f[(long)(N - 2) * N] - f[(long)(N - 3) * N];
            }
}

/* verlet_step, scp_sim.c:1583-1716: KDK with stored accelerations. */
void or_step(const OrParams *p, double *state, double *acc, int nsteps) {
    const int N = p->N; const long n3 = (long)N * N * N;
    const double hdt = 0.5 * p->dt, dt = p->dt;
    const int gauged = (p->mode == 2 && p->g != 0.0);
    const int nacc = or_nacc(p->mode);
    const int nkick = gauged ? 15 : (p->mode >= 1 ? 12 : 6);
    (void)nacc;
    for (int step = 0; step < nsteps; step++) {
        for (int b = 0; b < nkick; b++) {                      /* stage 1: C:1597-1613 */
            double *v = state + (long)acc_to_vel(b) * n3; const double *a = acc + (long)b * n3;
            #pragma omp parallel for schedule(static)
            for (long c = 0; c < n3; c++) v[c] += hdt * a[c];
        }
        for (int b = 0; b < (p->mode >= 1 ? 12 : 6); b++) {  /* stage 2: C:1629-1640 */
            int vc = acc_to_vel(b); double *x = state + (long)vel_to_pos(vc) * n3;
            const double *v = state + (long)vc * n3;
            #pragma omp parallel for schedule(static)
            for (long c = 0; c < n3; c++) x[c] += dt * v[c];
        }
        if (gauged) {                                          /* C:1641-1651 */
            const double gad = -p->g * p->dx * dt;
            for (int d = 0; d < 3; d++) {
                double *th = state + (24L + d) * n3; const double *E = state + (27L + d) * n3;
                #pragma omp parallel for schedule(static)
                for (long c = 0; c < n3; c++) {
                    double v = th[c] + gad * E[c];
                    th[c] = v - 2.0 * OR_PI * rint(v / (2.0 * OR_PI));
                }
            }
        }
        or_forces(p, state, acc);                              /* stage 3 */
        for (int b = 0; b < nkick; b++) {                      /* stage 4: C:1680-1696 */
            double *v = state + (long)acc_to_vel(b) * n3; const double *a = acc + (long)b * n3;
            #pragma omp parallel for schedule(static)
            for (long c = 0; c < n3; c++) v[c] += hdt * a[c];
        }
        if (p->bc_type == 0) damp(p, state);                  /* stage 5: C:1712-1715 */
        else if (p->bc_type == 1) gradient_bc(p, state, acc);
    }
}

/* ---------------- diagnostics ---------------- */

static void energy_real(const OrParams *p, const View *w, double *o) {
    const int N = p->N; const long n3 = (long)N * N * N;
    const double dx = p->dx, dV = dx * dx * dx, h1 = 1.0 / (2.0 * dx);
    const double MU = p->mu, KAPPA = p->kappa, MASS2 = p->m2, MT2 = p->mtheta2, ETA = p->eta;
    const int MODE = p->rmode; const double KG = p->kappa_gamma;
    double epk = 0, etk = 0, eg = 0, em = 0, ep = 0, etg = 0, etm = 0, ec = 0, pm = 0, Pm = 0;
    for (long c = 0; c < n3; c++) {                            /* scp_sim.c:1734-1796 */
        long np[3], nm[3]; neighbours(N, c, np, nm);
        double p0 = w->u[0][c], p1 = w->u[1][c], p2 = w->u[2][c];
        double sig = p0 * p0 + p1 * p1 + p2 * p2;
        double me2 = (MODE == 1) ? p->inv_alpha / (1.0 + p->inv_beta * sig) : MASS2;
        double keff = (MODE == 3) ? KAPPA / (1.0 + KG * sig) : KAPPA;
        for (int a = 0; a < 3; a++) {
            epk += 0.5 * w->ud[a][c] * w->ud[a][c] * dV;
            etk += 0.5 * w->tud[a][c] * w->tud[a][c] * dV;
            double gx = (w->u[a][np[0]] - w->u[a][nm[0]]) * h1;
            double gy = (w->u[a][np[1]] - w->u[a][nm[1]]) * h1;
            double gz = (w->u[a][np[2]] - w->u[a][nm[2]]) * h1;
            eg += 0.5 * (gx * gx + gy * gy + gz * gz) * dV;
            em += 0.5 * me2 * w->u[a][c] * w->u[a][c] * dV;
            double tx = (w->tu[a][np[0]] - w->tu[a][nm[0]]) * h1;
            double ty = (w->tu[a][np[1]] - w->tu[a][nm[1]]) * h1;
            double tz = (w->tu[a][np[2]] - w->tu[a][nm[2]]) * h1;
            etg += 0.5 * (tx * tx + ty * ty + tz * tz) * dV;
            etm += 0.5 * MT2 * w->tu[a][c] * w->tu[a][c] * dV;
            double ap = fabs(w->u[a][c]); if (ap > pm) pm = ap;
        }
        double P = p0 * p1 * p2, P2 = P * P;
        if (MODE == 3) {
            double D = 1.0 + KG * sig + KAPPA * P2;
            ep += (MU / 2.0) * P2 * (1.0 + KG * sig) / D * dV;
        } else {
            ep += (MU / 2.0) * P2 / (1.0 + keff * P2) * dV;
        }
        double Pa = fabs(P); if (Pa > Pm) Pm = Pa;
        if (p->alpha_cs != 0 || p->beta_h != 0 || p->kappa_h != 0) {
            double cx[3];
            for (int a = 0; a < 3; a++) cx[a] = curl_at(w->u, a, np, nm, h1);
            if (p->alpha_cs != 0) {
                double m0 = cx[0] * 0.5 - w->tu[0][c], m1 = cx[1] * 0.5 - w->tu[1][c], m2 = cx[2] * 0.5 - w->tu[2][c];
                ep += p->alpha_cs * (m0 * m0 + m1 * m1 + m2 * m2) * dV;
            }
            if (p->beta_h != 0) {
                double csq = cx[0] * cx[0] + cx[1] * cx[1] + cx[2] * cx[2];
                double T2 = w->tu[0][c] * w->tu[0][c] + w->tu[1][c] * w->tu[1][c] + w->tu[2][c] * w->tu[2][c];
                ep += 0.5 * p->beta_h * T2 * csq * dV;
            }
            if (p->kappa_h != 0) {
                double hel = p0 * cx[0] + p1 * cx[1] + p2 * cx[2];
                ep += p->kappa_h * P2 * hel * dV;
            }
        }
        for (int a = 0; a < 3; a++) {
            double ct = curl_at(w->tu, a, np, nm, h1);
            ec -= ETA * w->u[a][c] * ct * dV;
        }
    }
    o[ORD_E_PHI_KIN] = epk; o[ORD_E_THETA_KIN] = etk; o[ORD_E_GRAD] = eg; o[ORD_E_MASS] = em;
    o[ORD_E_POT] = ep; o[ORD_E_TGRAD] = etg; o[ORD_E_TMASS] = etm; o[ORD_E_COUPLING] = ec;
    o[ORD_E_TOTAL] = epk + etk + eg + em + ep + etg + etm + ec;
    o[ORD_PHI_MAX] = pm; o[ORD_P_MAX] = Pm;
    double ti = 0, pint = 0;
    for (long c = 0; c < n3; c++) {                            /* C:1803-1816 */
        for (int a = 0; a < 3; a++) ti += w->tu[a][c] * w->tu[a][c];
        pint += fabs(w->u[0][c] * w->u[1][c] * w->u[2][c]) * dV;
    }
    o[ORD_THETA_RMS] = sqrt(ti / (3.0 * n3));
    o[ORD_P_INT] = pint;
}

/* complex (gauged=0) and gauged energies: scp_sim.c:1823-1880 and 1886-2036 */
static void energy_complex(const OrParams *p, const View *w, double *o, int gauged) {
    const int N = p->N; const long n3 = (long)N * N * N;
    const double dx = p->dx, dV = dx * dx * dx, h1 = 1.0 / (2.0 * dx);
    const double MU = p->mu, KAPPA = p->kappa, MASS2 = p->m2, MT2 = p->mtheta2, ETA = p->eta;
    const double GG = p->g;
    const double BPRE = gauged ? 1.0 / (GG * GG * dx * dx * dx * dx) : 0.0;
    static const int C1[3] = {1, 2, 0}, C2[3] = {2, 0, 1};
    double epk = 0, etk = 0, eg = 0, em = 0, ep = 0, etg = 0, etm = 0, ec = 0, pm = 0, Pm = 0, eem = 0;
    for (long c = 0; c < n3; c++) {
        long np[3], nm[3]; neighbours(N, c, np, nm);
        double fu[6], fv[6];
        for (int a = 0; a < 3; a++) {
            fu[a] = w->u[a][c]; fv[a] = w->v[a][c]; fu[3 + a] = w->tu[a][c]; fv[3 + a] = w->tv[a][c];
        }
        double DU[3][6], DV[3][6];
        for (int d = 0; d < 3; d++) {
            if (gauged) {
                double cP = cos(w->th[d][c]), sP = sin(w->th[d][c]);
                double cM = cos(w->th[d][nm[d]]), sM = sin(w->th[d][nm[d]]);
                for (int a = 0; a < 3; a++) {
                    double upn = w->u[a][np[d]], vpn = w->v[a][np[d]], umn = w->u[a][nm[d]], vmn = w->v[a][nm[d]];
                    double tpn = w->tu[a][np[d]], wpn = w->tv[a][np[d]], tmn = w->tu[a][nm[d]], wmn = w->tv[a][nm[d]];
                    DU[d][a]     = ((cP * upn - sP * vpn) - (cM * umn + sM * vmn)) * h1;
                    DV[d][a]     = ((cP * vpn + sP * upn) - (cM * vmn - sM * umn)) * h1;
                    DU[d][3 + a] = ((cP * tpn - sP * wpn) - (cM * tmn + sM * wmn)) * h1;
                    DV[d][3 + a] = ((cP * wpn + sP * tpn) - (cM * wmn - sM * tmn)) * h1;
                }
            } else {
                for (int a = 0; a < 3; a++) {
                    DU[d][a]     = (w->u[a][np[d]] - w->u[a][nm[d]]) * h1;
                    DV[d][a]     = (w->v[a][np[d]] - w->v[a][nm[d]]) * h1;
                    DU[d][3 + a] = (w->tu[a][np[d]] - w->tu[a][nm[d]]) * h1;
                    DV[d][3 + a] = (w->tv[a][np[d]] - w->tv[a][nm[d]]) * h1;
                }
            }
        }
        double s2[3];
        for (int a = 0; a < 3; a++) {
            s2[a] = fu[a] * fu[a] + fv[a] * fv[a];
            epk += 0.5 * (w->ud[a][c] * w->ud[a][c] + w->vd[a][c] * w->vd[a][c]) * dV;
            etk += 0.5 * (w->tud[a][c] * w->tud[a][c] + w->tvd[a][c] * w->tvd[a][c]) * dV;
            if (gauged) {
                eg += 0.5 * (DU[0][a] * DU[0][a] + DU[1][a] * DU[1][a] + DU[2][a] * DU[2][a]
                           + DV[0][a] * DV[0][a] + DV[1][a] * DV[1][a] + DV[2][a] * DV[2][a]) * dV;
                em += 0.5 * MASS2 * s2[a] * dV;
                etg += 0.5 * (DU[0][3 + a] * DU[0][3 + a] + DU[1][3 + a] * DU[1][3 + a] + DU[2][3 + a] * DU[2][3 + a]
                            + DV[0][3 + a] * DV[0][3 + a] + DV[1][3 + a] * DV[1][3 + a] + DV[2][3 + a] * DV[2][3 + a]) * dV;
                etm += 0.5 * MT2 * (fu[3 + a] * fu[3 + a] + fv[3 + a] * fv[3 + a]) * dV;
            } else {
                /* C:1855-1864: gradient sums ordered gx,gy,gz,hx,hy,hz */
                eg += 0.5 * (DU[0][a] * DU[0][a] + DU[1][a] * DU[1][a] + DU[2][a] * DU[2][a]
                           + DV[0][a] * DV[0][a] + DV[1][a] * DV[1][a] + DV[2][a] * DV[2][a]) * dV;
                em += 0.5 * MASS2 * (fu[a] * fu[a] + fv[a] * fv[a]) * dV;
                etg += 0.5 * (DU[0][3 + a] * DU[0][3 + a] + DU[1][3 + a] * DU[1][3 + a] + DU[2][3 + a] * DU[2][3 + a]
                            + DV[0][3 + a] * DV[0][3 + a] + DV[1][3 + a] * DV[1][3 + a] + DV[2][3 + a] * DV[2][3 + a]) * dV;
                etm += 0.5 * MT2 * (fu[3 + a] * fu[3 + a] + fv[3 + a] * fv[3 + a]) * dV;
            }
            double mod = sqrt(s2[a]); if (mod > pm) pm = mod;
            double reDxT, imDxT;
            if (gauged) {
                reDxT = DU[C1[a]][3 + C2[a]] - DU[C2[a]][3 + C1[a]];
                imDxT = DV[C1[a]][3 + C2[a]] - DV[C2[a]][3 + C1[a]];
            } else {
                const double *tu3[3] = {w->tu[0], w->tu[1], w->tu[2]}, *tv3[3] = {w->tv[0], w->tv[1], w->tv[2]};
                reDxT = curl_at(tu3, a, np, nm, h1);
                imDxT = curl_at(tv3, a, np, nm, h1);
            }
            ec -= ETA * (fu[a] * reDxT + fv[a] * imDxT) * dV;
        }
        double s = s2[0] * s2[1] * s2[2];
        ep += (MU / 2.0) * s / (1.0 + KAPPA * s) * dV;
        double Pmod = sqrt(s); if (Pmod > Pm) Pm = Pmod;
        if (gauged) {
            double ee = 0.5 * (w->E[0][c] * w->E[0][c] + w->E[1][c] * w->E[1][c] + w->E[2][c] * w->E[2][c]);
            static const int PA[3] = {0, 1, 2}, PB[3] = {1, 2, 0};
            double bb = 0.0;
            for (int pl = 0; pl < 3; pl++) {
                int a = PA[pl], b = PB[pl];
                double ang = w->th[a][c] + w->th[b][np[a]] - w->th[a][np[b]] - w->th[b][c];
                bb += 1.0 - cos(ang);
            }
            eem += (ee + BPRE * bb) * dV;
        }
    }
    o[ORD_E_PHI_KIN] = epk; o[ORD_E_THETA_KIN] = etk; o[ORD_E_GRAD] = eg; o[ORD_E_MASS] = em;
    o[ORD_E_POT] = ep; o[ORD_E_TGRAD] = etg; o[ORD_E_TMASS] = etm; o[ORD_E_COUPLING] = ec;
    o[ORD_E_TOTAL] = epk + etk + eg + em + ep + etg + etm + ec + (gauged ? eem : 0.0);
    o[ORD_PHI_MAX] = pm; o[ORD_P_MAX] = Pm;
    double ti = 0, pint = 0;
    for (long c = 0; c < n3; c++) {                            /* C:2039-2059 */
        for (int a = 0; a < 3; a++) ti += w->tu[a][c] * w->tu[a][c] + w->tv[a][c] * w->tv[a][c];
        double a0 = w->u[0][c] * w->u[0][c] + w->v[0][c] * w->v[0][c];
        double a1 = w->u[1][c] * w->u[1][c] + w->v[1][c] * w->v[1][c];
        double a2 = w->u[2][c] * w->u[2][c] + w->v[2][c] * w->v[2][c];
        pint += sqrt(a0 * a1 * a2) * dV;
    }
    o[ORD_THETA_RMS] = sqrt(ti / (6.0 * n3));
    o[ORD_P_INT] = pint;
    o[ORD_E_EM] = eem;
}

/* compute_charges, scp_sim.c:2074-2169 */
static void charges(const OrParams *p, const View *w, double *o) {
    const int N = p->N; const long n3 = (long)N * N * N, NN = (long)N * N;
    const double L = p->L, dx = p->dx, dV = dx * dx * dx;
    const double R2 = p->qdiag_radius * p->qdiag_radius;
    double qp = 0, qt = 0, smax = 0, M0 = 0, cx = 0, cy = 0, cz = 0, qc = 0, qa[3] = {0, 0, 0};
    for (long c = 0; c < n3; c++) {
        int i = (int)(c / NN), j = (int)((c / N) % N), k = (int)(c % N);
        double x = -L + i * dx, y = -L + j * dx, z = -L + k * dx;
        double rho2 = 0, s = 1.0, qpv = 0, qtv = 0;
        for (int a = 0; a < 3; a++) {
            double u = w->u[a][c], v = w->v[a][c], ud = w->ud[a][c], vd = w->vd[a][c];
            double tu = w->tu[a][c], tv = w->tv[a][c], tud = w->tud[a][c], tvd = w->tvd[a][c];
            double q = u * vd - v * ud;
            qpv += q; qa[a] += q;
            qtv += tu * tvd - tv * tud;
            double s2a = u * u + v * v;
            rho2 += s2a; s *= s2a;
        }
        qp += qpv; qt += qtv;
        if (x * x + y * y + z * z <= R2) qc += qpv + qtv;
        if (s > smax) smax = s;
        M0 += rho2; cx += x * rho2; cy += y * rho2; cz += z * rho2;
    }
    o[ORD_Q_PHI] = qp * dV; o[ORD_Q_THETA] = qt * dV; o[ORD_Q_TOTAL] = qp * dV + qt * dV;
    o[ORD_S_MAX] = smax; o[ORD_Q_CORE] = qc * dV;
    for (int a = 0; a < 3; a++) o[ORD_Q_P0 + a] = qa[a] * dV;
    long arg = -1; const double thr = smax * (1.0 - 1e-12);
    for (long c = 0; c < n3 && arg < 0; c++) {
        double s = 1.0;
        for (int a = 0; a < 3; a++) s *= w->u[a][c] * w->u[a][c] + w->v[a][c] * w->v[a][c];
        if (s >= thr) arg = c;
    }
    o[ORD_OMEGA_CORE] = 0.0;
    if (arg >= 0) {
        double u0 = w->u[0][arg], v0 = w->v[0][arg], ud0 = w->ud[0][arg], vd0 = w->vd[0][arg];
        double den = u0 * u0 + v0 * v0;
        if (den >= 1e-20) o[ORD_OMEGA_CORE] = (u0 * vd0 - v0 * ud0) / den;
    }
    int jc = (int)lround(L / dx);                    if (jc < 0) jc = 0; if (jc > N - 1) jc = N - 1;
    int i1 = (int)lround((p->qdiag_probe1 + L) / dx); if (i1 < 0) i1 = 0; if (i1 > N - 1) i1 = N - 1;
    int i2 = (int)lround((p->qdiag_probe2 + L) / dx); if (i2 < 0) i2 = 0; if (i2 > N - 1) i2 = N - 1;
    long p1 = (long)i1 * NN + (long)jc * N + jc, p2 = (long)i2 * NN + (long)jc * N + jc;
    o[ORD_THP1U] = w->tu[1][p1]; o[ORD_THP1V] = w->tv[1][p1];
    o[ORD_THP2U] = w->tu[1][p2]; o[ORD_THP2V] = w->tv[1][p2];
    M0 *= dV;
    if (M0 < 1e-30) { o[ORD_R_CORE] = 0; return; }
    cx = cx * dV / M0; cy = cy * dV / M0; cz = cz * dV / M0;
    double rs = 0;
    for (long c = 0; c < n3; c++) {
        int i = (int)(c / NN), j = (int)((c / N) % N), k = (int)(c % N);
        double x = -L + i * dx, y = -L + j * dx, z = -L + k * dx;
        double rho2 = 0;
        for (int a = 0; a < 3; a++) rho2 += w->u[a][c] * w->u[a][c] + w->v[a][c] * w->v[a][c];
        rs += ((x - cx) * (x - cx) + (y - cy) * (y - cy) + (z - cz) * (z - cz)) * rho2;
    }
    o[ORD_R_CORE] = sqrt(rs * dV / M0);
}

static inline double rho_at(const View *w, long c) {   /* fabric_rho_at, C:174-180 */
    double r = 0;
    for (int a = 0; a < 3; a++)
        r += w->u[a][c] * w->vd[a][c] - w->v[a][c] * w->ud[a][c]
           + w->tu[a][c] * w->tvd[a][c] - w->tv[a][c] * w->tud[a][c];
    return r;
}

static inline double gauss_at(const OrParams *p, const View *w, long c, const long nm[3]) {  /* C:2176-2181 */
    double divE = (w->E[0][c] - w->E[0][nm[0]] + w->E[1][c] - w->E[1][nm[1]]
                 + w->E[2][c] - w->E[2][nm[2]]) / p->dx;
    return divE - p->g * rho_at(w, c);
}

/* compute_gauss, scp_sim.c:2186-2243 */
static void gauss_diag(const OrParams *p, const View *w, double *o) {
    const int N = p->N; const long n3 = (long)N * N * N, NN = (long)N * N;
    const double L = p->L, dx = p->dx, dV = dx * dx * dx, GG = p->g, OFF = p->G_offset;
    const double Rint = (p->bc_type == 0) ? (L - p->damp_width - dx) : 1e300;
    const double Rint2 = Rint * Rint, Rc = p->qdiag_radius;
    double gmax = 0, gl2 = 0, eem = 0, fsum = 0; long nO = 0, nC = 0;
    for (long c = 0; c < n3; c++) {
        int i = (int)(c / NN), j = (int)((c / N) % N), k = (int)(c % N);
        long np[3], nm[3]; neighbours(N, c, np, nm);
        double x = -L + i * dx, y = -L + j * dx, z = -L + k * dx;
        double divE = (w->E[0][c] - w->E[0][nm[0]] + w->E[1][c] - w->E[1][nm[1]] + w->E[2][c] - w->E[2][nm[2]]) / dx;
        double G = gauss_at(p, w, c, nm);
        if (x * x + y * y + z * z < Rint2) {
            double d = fabs(G - OFF); if (d > gmax) gmax = d;
            gl2 += (G - OFF) * (G - OFF); nO++;
        }
        double ee = 0.5 * (w->E[0][c] * w->E[0][c] + w->E[1][c] * w->E[1][c] + w->E[2][c] * w->E[2][c]);
        if (GG != 0) {
            static const int PA[3] = {0, 1, 2}, PB[3] = {1, 2, 0};
            double bb = 0;
            for (int pl = 0; pl < 3; pl++) {
                int a = PA[pl], b = PB[pl];
                double ang = w->th[a][c] + w->th[b][np[a]] - w->th[a][np[b]] - w->th[b][c];
                bb += 1.0 - cos(ang);
            }
            ee += bb / (GG * GG * dx * dx * dx * dx);
        }
        eem += ee * dV;
        if (fabs(x) <= Rc && fabs(y) <= Rc && fabs(z) <= Rc) { fsum += divE * dV; nC++; }
    }
    o[ORD_GAUSS_MAX] = gmax;
    o[ORD_GAUSS_L2] = nO > 0 ? sqrt(gl2 / (double)nO) : 0.0;
    o[ORD_E_EM] = eem;
    o[ORD_Q_FLUX] = (GG != 0) ? (fsum - OFF * (double)nC * dV) / GG : 0.0;
}

void or_diag(const OrParams *p, const double *state, double *o) {
    const long n3 = (long)p->N * p->N * p->N;
    for (int q = 0; q < ORD_COUNT; q++) o[q] = 0.0;
    View w = make_view(state, n3, p->mode);
    if (p->mode == 0) { energy_real(p, &w, o); return; }
    const int gauged = (p->mode == 2 && p->g != 0.0);
    energy_complex(p, &w, o, gauged);
    charges(p, &w, o);
    if (p->mode == 2) {
        double eem_energy = o[ORD_E_EM];
        gauss_diag(p, &w, o);
        (void)eem_energy;   /* diag.tsv's E_em column comes from compute_gauss (C:3048-3049) */
    }
}

/* net_charge_check, scp_sim.c:2264-2279. Returns Q_net, writes Q_abs. */
double or_net_charge(const OrParams *p, const double *state, double *q_abs) {
    const long n3 = (long)p->N * p->N * p->N;
    const double dV = p->dx * p->dx * p->dx;
    View w = make_view(state, n3, 2);
    double qn = 0, qa = 0;
    for (long c = 0; c < n3; c++) { double r = rho_at(&w, c); qn += r * dV; qa += fabs(r) * dV; }
    if (q_abs) *q_abs = qa;
    return qn;
}

/* init_gauss_project, scp_sim.c:2284-2400: CG on the periodic 7-point Poisson
 * problem (-lap) chi = G - Gbar, E += forward gradient of chi, freeze G_offset.
 * Returns the iteration count (-1 if CG failed). */
int or_gauss_project(OrParams *p, double *state, double *gauss_max0) {
    const int N = p->N; const long n3 = (long)N * N * N;
    const double dx = p->dx, h2 = 1.0 / (dx * dx);
    View w = make_view(state, n3, 2);
    double *b = malloc(n3 * sizeof(double)), *x = calloc(n3, sizeof(double));
    double *r = malloc(n3 * sizeof(double)), *q = malloc(n3 * sizeof(double)), *Aq = malloc(n3 * sizeof(double));
    double gsum = 0, gmax0 = 0;
    for (long c = 0; c < n3; c++) { long np[3], nm[3]; neighbours(N, c, np, nm); b[c] = gauss_at(p, &w, c, nm); gsum += b[c]; }
    double gbar = gsum / (double)n3;
    for (long c = 0; c < n3; c++) { b[c] -= gbar; double d = fabs(b[c]); if (d > gmax0) gmax0 = d; }
    const double tol = 1e-13 * fmax(1.0, gmax0);
    memcpy(r, b, n3 * sizeof(double)); memcpy(q, b, n3 * sizeof(double));
    double rz = 0; for (long c = 0; c < n3; c++) rz += r[c] * r[c];
    int it = 0; double rmax = gmax0;
    while (rmax > tol && it < 20000) {
        double qAq = 0;
        for (long c = 0; c < n3; c++) {
            long np[3], nm[3]; neighbours(N, c, np, nm);
            double lap = (q[np[0]] + q[nm[0]] + q[np[1]] + q[nm[1]] + q[np[2]] + q[nm[2]] - 6.0 * q[c]) * h2;
            Aq[c] = -lap; qAq += q[c] * Aq[c];
        }
        double alpha = rz / qAq, rz2 = 0, xs = 0, rm = 0;
        for (long c = 0; c < n3; c++) {
            x[c] += alpha * q[c]; r[c] -= alpha * Aq[c];
            rz2 += r[c] * r[c]; xs += x[c];
            double d = fabs(r[c]); if (d > rm) rm = d;
        }
        double xbar = xs / (double)n3, beta = rz2 / rz;
        for (long c = 0; c < n3; c++) { x[c] -= xbar; q[c] = r[c] + beta * q[c]; }
        rz = rz2; rmax = rm; it++;
    }
    int ok = rmax <= tol;
    if (ok) {
        double *E[3] = {state + 27 * n3, state + 28 * n3, state + 29 * n3};
        for (long c = 0; c < n3; c++) {
            long np[3], nm[3]; neighbours(N, c, np, nm);
            for (int d = 0; d < 3; d++) E[d][c] += (x[np[d]] - x[c]) / dx;
        }
        double os = 0;
        for (long c = 0; c < n3; c++) { long np[3], nm[3]; neighbours(N, c, np, nm); os += gauss_at(p, &w, c, nm); }
        p->G_offset = os / (double)n3;
    }
    if (gauss_max0) *gauss_max0 = gmax0;
    free(b); free(x); free(r); free(q); free(Aq);
    return ok ? it : -1;
}

/* The random lattice gauge transformation of do_test_gauge_xform,
 * scp_sim.c:2485-2514: matter rotates by e^{-i g alpha(x)}, links shift by
 * g (alpha(x+i) - alpha(x)), alpha from a counter-keyed LCG. */
void or_gauge_transform(const OrParams *p, double *state) {
    const int N = p->N; const long n3 = (long)N * N * N;
    const double GG = p->g;
    double *al = malloc(n3 * sizeof(double));
    for (long c = 0; c < n3; c++) {
        uint64_t s = (uint64_t)c * 6364136223846793005ULL + 1442695040888963407ULL;
        s = s * 6364136223846793005ULL + 1442695040888963407ULL;
        al[c] = (((double)(s >> 11)) / 9007199254740992.0 - 0.5) * 20.0;
    }
    static const int PAIRS[4][2] = {{0, 12}, {6, 18}, {3, 15}, {9, 21}};  /* (re, im) column bases */
    for (long c = 0; c < n3; c++) {
        long np[3], nm[3]; neighbours(N, c, np, nm);
        double ca = cos(GG * al[c]), sa = sin(GG * al[c]);
        for (int pr = 0; pr < 4; pr++)
            for (int a = 0; a < 3; a++) {
                double *re = state + (long)(PAIRS[pr][0] + a) * n3 + c, *im = state + (long)(PAIRS[pr][1] + a) * n3 + c;
                double u = *re, v = *im;
                *re = ca * u + sa * v; *im = ca * v - sa * u;
            }
        for (int d = 0; d < 3; d++) {
            double *th = state + (24L + d) * n3;
            double v = th[c] + GG * (al[np[d]] - al[c]);
            th[c] = v - 2.0 * OR_PI * rint(v / (2.0 * OR_PI));
        }
    }
    free(al);
}


/* ------------------------------------------------------------------------ */
/* fixed-charge gradient flow, sfa/seed/eta_qflow.c (plain fixed-Q mode)     */
/* ------------------------------------------------------------------------ */

double or_qflow_norm(const OrQflow *p, const double *st) {
    const long N3 = (long)p->N * p->N * p->N;
    const double dx = 2.0 * p->L / (p->N - 1), dV = dx * dx * dx;
    const double *u = st, *v = st + 3 * N3, *tu = st + 6 * N3, *tv = st + 9 * N3;
    double Nn = 0;
    for (long ii = 0; ii < N3; ii++)                       /* eta_qflow.c:193-197 */
        for (int a = 0; a < 3; a++) {
            const long q = a * N3 + ii;
            Nn += u[q] * u[q] + v[q] * v[q] + tu[q] * tu[q] + tv[q] * tv[q];
        }
    return Nn * dV;
}

int or_qflow(const OrQflow *p, double *st, double Q, double dtau, int iters, int printev, double *rows,
             int maxrows, double *omega_last) {
    const int N = p->N;
    const long N3 = (long)N * N * N, NN = (long)N * N;
    const double L = p->L, dx = 2.0 * L / (N - 1), dx2 = dx * dx, dV = dx * dx * dx;
    const double invdx2 = 1.0 / dx2, inv2dx = 1.0 / (2.0 * dx);
    const double eta = p->eta, m = p->m, mth = p->mth, mu = p->mu, kappa = p->kappa;
    double *nx = malloc(sizeof(double) * 12 * (size_t)N3);
    double *cur = st, *nxt = nx;
    int nrows = 0;
    double omega = 0;
#define WR(a) (((a) + N) % N)
#define IX3(i, j, k) (((long)(i) * N + (j)) * N + (k))
    for (int it = 0; it <= iters; it++) {
        const double *u = cur, *v = cur + 3 * N3, *tu = cur + 6 * N3, *tv = cur + 9 * N3;
        double Nn = 0;                                      /* eta_qflow.c:269-274 */
        for (long ii = 0; ii < N3; ii++)
            for (int a = 0; a < 3; a++) {
                const long q = a * N3 + ii;
                Nn += u[q] * u[q] + v[q] * v[q] + tu[q] * tu[q] + tv[q] * tv[q];
            }
        Nn *= dV;
        omega = Q / Nn;
        const double m2w = m * m - omega * omega, mt2w = mth * mth - omega * omega;
        double maxF = 0;
        for (long ii = 0; ii < N3; ii++) {                 /* eta_qflow.c:278-330 */
            const int i = (int)(ii / NN), j = (int)((ii / N) % N), k = (int)(ii % N);
            const long X[2][3] = {{IX3(WR(i + 1), j, k), IX3(i, WR(j + 1), k), IX3(i, j, WR(k + 1))},
                                  {IX3(WR(i - 1), j, k), IX3(i, WR(j - 1), k), IX3(i, j, WR(k - 1))}};
            const double r0 = u[ii] * u[ii] + v[ii] * v[ii];
            const double r1 = u[N3 + ii] * u[N3 + ii] + v[N3 + ii] * v[N3 + ii];
            const double r2 = u[2 * N3 + ii] * u[2 * N3 + ii] + v[2 * N3 + ii] * v[2 * N3 + ii];
            const double s = r0 * r1 * r2, den = 1.0 + kappa * s;
            const double Vtp = 0.5 * mu / (den * den);
            const double Pa[3] = {r1 * r2, r0 * r2, r0 * r1};
            for (int a = 0; a < 3; a++) {
                const double *ua = u + a * N3, *va = v + a * N3, *ta = tu + a * N3, *sa = tv + a * N3;
#define LAP(f) (((f)[X[0][0]] + (f)[X[1][0]] + (f)[X[0][1]] + (f)[X[1][1]] + (f)[X[0][2]] + (f)[X[1][2]] - 6.0 * (f)[ii]) * invdx2)
#define DRV(f, ax, comp) ((f)[(comp) * N3 + X[0][ax]] - (f)[(comp) * N3 + X[1][ax]])
                const double lu = LAP(ua), lv = LAP(va), ltu = LAP(ta), ltv = LAP(sa);
                const int b = (a + 1) % 3, c = (a + 2) % 3;
                const double curl_tu = (DRV(tu, b, c) - DRV(tu, c, b)) * inv2dx;
                const double curl_tv = (DRV(tv, b, c) - DRV(tv, c, b)) * inv2dx;
                const double curl_u = (DRV(u, b, c) - DRV(u, c, b)) * inv2dx;
                const double curl_v = (DRV(v, b, c) - DRV(v, c, b)) * inv2dx;
                const double Fu = lu - m2w * ua[ii] - 2.0 * Vtp * ua[ii] * Pa[a] + eta * curl_tu;
                const double Fv = lv - m2w * va[ii] - 2.0 * Vtp * va[ii] * Pa[a] + eta * curl_tv;
                const double Ftu = ltu - mt2w * ta[ii] + eta * curl_u;
                const double Ftv = ltv - mt2w * sa[ii] + eta * curl_v;
                nxt[a * N3 + ii] = ua[ii] + dtau * Fu;
                nxt[(3 + a) * N3 + ii] = va[ii] + dtau * Fv;
                nxt[(6 + a) * N3 + ii] = ta[ii] + dtau * Ftu;
                nxt[(9 + a) * N3 + ii] = sa[ii] + dtau * Ftv;
                double af = fabs(Fu);
                if (af > maxF) maxF = af;
                af = fabs(Ftu);
                if (af > maxF) maxF = af;
            }
        }
        double *t = cur; cur = nxt; nxt = t;
        if (it % printev == 0 || it == iters) {               /* eta_qflow.c:336-373 */
            const double *U = cur, *V = cur + 3 * N3, *TU = cur + 6 * N3, *TV = cur + 9 * N3;
            double Eg = 0, Em = 0, Ep = 0, Ex = 0, smax = 0;
            for (long ii = 0; ii < N3; ii++) {
                const int i = (int)(ii / NN), j = (int)((ii / N) % N), k = (int)(ii % N);
                const long X[2][3] = {{IX3(WR(i + 1), j, k), IX3(i, WR(j + 1), k), IX3(i, j, WR(k + 1))},
                                      {IX3(WR(i - 1), j, k), IX3(i, WR(j - 1), k), IX3(i, j, WR(k - 1))}};
                const double r0 = U[ii] * U[ii] + V[ii] * V[ii];
                const double r1 = U[N3 + ii] * U[N3 + ii] + V[N3 + ii] * V[N3 + ii];
                const double r2 = U[2 * N3 + ii] * U[2 * N3 + ii] + V[2 * N3 + ii] * V[2 * N3 + ii];
                const double s = r0 * r1 * r2;
                if (s > smax) smax = s;
                Ep += 0.5 * mu * s / (1.0 + kappa * s);
                for (int a = 0; a < 3; a++) {
                    const double *F4[4] = {U + a * N3, V + a * N3, TU + a * N3, TV + a * N3};
                    for (int q = 0; q < 4; q++) {
                        const double *f = F4[q];
                        const double gx = (f[X[0][0]] - f[X[1][0]]) * inv2dx, gy = (f[X[0][1]] - f[X[1][1]]) * inv2dx;
                        const double gz = (f[X[0][2]] - f[X[1][2]]) * inv2dx;
                        Eg += 0.5 * (gx * gx + gy * gy + gz * gz);
                    }
                    Em += 0.5 * m * m * (U[a * N3 + ii] * U[a * N3 + ii] + V[a * N3 + ii] * V[a * N3 + ii])
                        + 0.5 * mth * mth * (TU[a * N3 + ii] * TU[a * N3 + ii] + TV[a * N3 + ii] * TV[a * N3 + ii]);
                    const int b = (a + 1) % 3, c = (a + 2) % 3;
                    const double curl_u = (DRV(U, b, c) - DRV(U, c, b)) * inv2dx;
                    const double curl_v = (DRV(V, b, c) - DRV(V, c, b)) * inv2dx;
                    Ex += -eta * (TU[a * N3 + ii] * curl_u + TV[a * N3 + ii] * curl_v);
                }
            }
            const double EQ = (Eg + Em + Ep + Ex) * dV + Q * Q / (2.0 * Nn);
            if (nrows < maxrows) {
                double *r = rows + 6 * nrows;
                r[0] = it; r[1] = omega; r[2] = EQ; r[3] = maxF; r[4] = smax; r[5] = Nn;
            }
            nrows++;
        }
#undef LAP
#undef DRV
    }
#undef WR
#undef IX3
    if (cur != st) memcpy(st, cur, sizeof(double) * 12 * (size_t)N3);
    free(nx);
    if (omega_last) *omega_last = omega;
    return nrows;
}

double or_qflow_wmask(const OrQflow *p, int i, int j) {
    const double dx = 2.0 * p->L / (p->N - 1);
    const double x = -p->L + i * dx, y = -p->L + j * dx, rho = sqrt(x * x + y * y);
    return 1.0 / (1.0 + exp((rho - 11.0) / 0.5));           /* eta_qflow.c:215-216 */
}

/* Jhat of the 12 fields of `st` into jh, with the sums N, ell, K in site order
 * (eta_qflow.c:232-252 and :381-396; the same expressions, the sums as one thread
 * of the reference accumulates them) */
static void qflow_jhat(const OrQflow *p, const double *st, double *jh, double *Nn, double *el, double *kk) {
    const int N = p->N;
    const long N3 = (long)N * N * N, NN = (long)N * N;
    const double L = p->L, dx = 2.0 * L / (N - 1), inv2dx = 1.0 / (2.0 * dx);
    double sn = 0, se = 0, sk = 0;
#define WR(a) (((a) + N) % N)
#define IX3(i, j, k) (((long)(i) * N + (j)) * N + (k))
    for (long ii = 0; ii < N3; ii++) {
        const int i = (int)(ii / NN), j = (int)((ii / N) % N), k = (int)(ii % N);
        const long Xp = IX3(WR(i + 1), j, k), Xm = IX3(WR(i - 1), j, k);
        const long Yp = IX3(i, WR(j + 1), k), Ym = IX3(i, WR(j - 1), k);
        const double x = -L + i * dx, y = -L + j * dx;
        for (int a = 0; a < 3; a++) {
            double jw[4];
            for (int q = 0; q < 4; q++) {                      /* u, v, tu, tv */
                const double *w = st + (long)(3 * q) * N3;
                const double spin = a == 0 ? w[N3 + ii] : a == 1 ? -w[ii] : 0.0;
                const double *wa = w + a * N3;
                jw[q] = x * (wa[Yp] - wa[Ym]) * inv2dx - y * (wa[Xp] - wa[Xm]) * inv2dx + spin;
                jh[(long)(3 * q + a) * N3 + ii] = jw[q];
            }
            const double u = st[a * N3 + ii], v = st[(3 + a) * N3 + ii];
            const double tu = st[(6 + a) * N3 + ii], tv = st[(9 + a) * N3 + ii];
            if (Nn) sn += u * u + v * v + tu * tu + tv * tv;
            se += u * jw[1] - v * jw[0] + tu * jw[3] - tv * jw[2];
            sk += jw[0] * jw[0] + jw[1] * jw[1] + jw[2] * jw[2] + jw[3] * jw[3];
        }
    }
#undef WR
#undef IX3
    if (Nn) *Nn = sn;
    *el = se;
    *kk = sk;
}

int or_qflow_j(const OrQflow *p, double *st, double Q, double Jt, double dtau, int iters, int printev, double *rows,
               int maxrows, double *out, double *jhat_out) {
    const int N = p->N;
    const long N3 = (long)N * N * N, NN = (long)N * N;
    const double L = p->L, dx = 2.0 * L / (N - 1), dx2 = dx * dx, dV = dx * dx * dx;
    const double invdx2 = 1.0 / dx2, inv2dx = 1.0 / (2.0 * dx);
    const double eta = p->eta, m = p->m, mth = p->mth, mu = p->mu, kappa = p->kappa;
    const double OMEGA_CAP = 0.15;                            /* eta_qflow.c:208 */
    double *nx = malloc(sizeof(double) * 12 * (size_t)N3);
    double *jh = malloc(sizeof(double) * 12 * (size_t)N3);
    double *wmask = malloc(sizeof(double) * (size_t)NN);
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++) wmask[(long)i * N + j] = or_qflow_wmask(p, i, j);
    double *cur = st, *nxt = nx;
    int nrows = 0;
    double omega = 0, Omega = 0, ell = 0, Kk = 0, cond = 1.0;
#define WR(a) (((a) + N) % N)
#define IX3(i, j, k) (((long)(i) * N + (j)) * N + (k))
    for (int it = 0; it <= iters; it++) {
        const double *u = cur, *v = cur + 3 * N3, *tu = cur + 6 * N3, *tv = cur + 9 * N3;
        double Nn, el, kk;                                   /* eta_qflow.c:232-262 */
        qflow_jhat(p, cur, jh, &Nn, &el, &kk);
        Nn *= dV; ell = el * dV; Kk = kk * dV;
        const double D = Nn * Kk - ell * ell, Dr = D + 1e-10 * Nn * Kk;
        cond = D / (Nn * Kk);
        omega = (Kk * Q - ell * Jt) / Dr;
        Omega = (ell * Q - Nn * Jt) / Dr;
        if (Omega > OMEGA_CAP) { Omega = OMEGA_CAP; omega = (Q + Omega * ell) / Nn; }
        if (Omega < -OMEGA_CAP) { Omega = -OMEGA_CAP; omega = (Q + Omega * ell) / Nn; }
        const double m2w = m * m - omega * omega, mt2w = mth * mth - omega * omega;
        double maxF = 0;
        for (long ii = 0; ii < N3; ii++) {                 /* eta_qflow.c:268-318 */
            const int i = (int)(ii / NN), j = (int)((ii / N) % N), k = (int)(ii % N);
            const long X[2][3] = {{IX3(WR(i + 1), j, k), IX3(i, WR(j + 1), k), IX3(i, j, WR(k + 1))},
                                  {IX3(WR(i - 1), j, k), IX3(i, WR(j - 1), k), IX3(i, j, WR(k - 1))}};
            const double r0 = u[ii] * u[ii] + v[ii] * v[ii];
            const double r1 = u[N3 + ii] * u[N3 + ii] + v[N3 + ii] * v[N3 + ii];
            const double r2 = u[2 * N3 + ii] * u[2 * N3 + ii] + v[2 * N3 + ii] * v[2 * N3 + ii];
            const double s = r0 * r1 * r2, den = 1.0 + kappa * s;
            const double Vtp = 0.5 * mu / (den * den);
            const double Pa[3] = {r1 * r2, r0 * r2, r0 * r1};
            const double x = -L + i * dx, y = -L + j * dx, wm = wmask[(long)i * N + j];
            const double wO = 2.0 * omega * Omega * wm, O2 = Omega * Omega * wm;
            for (int a = 0; a < 3; a++) {
                const double *ua = u + a * N3, *va = v + a * N3, *ta = tu + a * N3, *sa = tv + a * N3;
#define LAP(f) (((f)[X[0][0]] + (f)[X[1][0]] + (f)[X[0][1]] + (f)[X[1][1]] + (f)[X[0][2]] + (f)[X[1][2]] - 6.0 * (f)[ii]) * invdx2)
#define DRV(f, ax, comp) ((f)[(comp) * N3 + X[0][ax]] - (f)[(comp) * N3 + X[1][ax]])
                const double lu = LAP(ua), lv = LAP(va), ltu = LAP(ta), ltv = LAP(sa);
                const int b = (a + 1) % 3, c = (a + 2) % 3;
                const double curl_tu = (DRV(tu, b, c) - DRV(tu, c, b)) * inv2dx;
                const double curl_tv = (DRV(tv, b, c) - DRV(tv, c, b)) * inv2dx;
                const double curl_u = (DRV(u, b, c) - DRV(u, c, b)) * inv2dx;
                const double curl_v = (DRV(v, b, c) - DRV(v, c, b)) * inv2dx;
                double Fu = lu - m2w * ua[ii] - 2.0 * Vtp * ua[ii] * Pa[a] + eta * curl_tu;
                double Fv = lv - m2w * va[ii] - 2.0 * Vtp * va[ii] * Pa[a] + eta * curl_tv;
                double Ftu = ltu - mt2w * ta[ii] + eta * curl_u;
                double Ftv = ltv - mt2w * sa[ii] + eta * curl_v;
                double j2[4];                                   /* Jhat(Jhat w) of u, v, tu, tv */
                for (int q = 0; q < 4; q++) {
                    const double *w = jh + (long)(3 * q) * N3, *wa = w + a * N3;
                    const double spin = a == 0 ? w[N3 + ii] : a == 1 ? -w[ii] : 0.0;
                    j2[q] = x * (wa[X[0][1]] - wa[X[1][1]]) * inv2dx - y * (wa[X[0][0]] - wa[X[1][0]]) * inv2dx + spin;
                }
                Fu += -wO * jh[(3 + a) * N3 + ii] - O2 * j2[0];
                Fv += +wO * jh[a * N3 + ii] - O2 * j2[1];
                Ftu += -wO * jh[(9 + a) * N3 + ii] - O2 * j2[2];
                Ftv += +wO * jh[(6 + a) * N3 + ii] - O2 * j2[3];
                nxt[a * N3 + ii] = ua[ii] + dtau * Fu;
                nxt[(3 + a) * N3 + ii] = va[ii] + dtau * Fv;
                nxt[(6 + a) * N3 + ii] = ta[ii] + dtau * Ftu;
                nxt[(9 + a) * N3 + ii] = sa[ii] + dtau * Ftv;
                double af = fabs(Fu);
                if (af > maxF) maxF = af;
                af = fabs(Ftu);
                if (af > maxF) maxF = af;
            }
        }
        double *t = cur; cur = nxt; nxt = t;
        if (it % printev == 0 || it == iters) {               /* eta_qflow.c:336-358 */
            const double *U = cur, *V = cur + 3 * N3, *TU = cur + 6 * N3, *TV = cur + 9 * N3;
            double Eg = 0, Em = 0, Ep = 0, Ex = 0, smax = 0;
            for (long ii = 0; ii < N3; ii++) {
                const int i = (int)(ii / NN), j = (int)((ii / N) % N), k = (int)(ii % N);
                const long X[2][3] = {{IX3(WR(i + 1), j, k), IX3(i, WR(j + 1), k), IX3(i, j, WR(k + 1))},
                                      {IX3(WR(i - 1), j, k), IX3(i, WR(j - 1), k), IX3(i, j, WR(k - 1))}};
                const double r0 = U[ii] * U[ii] + V[ii] * V[ii];
                const double r1 = U[N3 + ii] * U[N3 + ii] + V[N3 + ii] * V[N3 + ii];
                const double r2 = U[2 * N3 + ii] * U[2 * N3 + ii] + V[2 * N3 + ii] * V[2 * N3 + ii];
                const double s = r0 * r1 * r2;
                if (s > smax) smax = s;
                Ep += 0.5 * mu * s / (1.0 + kappa * s);
                for (int a = 0; a < 3; a++) {
                    const double *F4[4] = {U + a * N3, V + a * N3, TU + a * N3, TV + a * N3};
                    for (int q = 0; q < 4; q++) {
                        const double *f = F4[q];
                        const double gx = (f[X[0][0]] - f[X[1][0]]) * inv2dx, gy = (f[X[0][1]] - f[X[1][1]]) * inv2dx;
                        const double gz = (f[X[0][2]] - f[X[1][2]]) * inv2dx;
                        Eg += 0.5 * (gx * gx + gy * gy + gz * gz);
                    }
                    Em += 0.5 * m * m * (U[a * N3 + ii] * U[a * N3 + ii] + V[a * N3 + ii] * V[a * N3 + ii])
                        + 0.5 * mth * mth * (TU[a * N3 + ii] * TU[a * N3 + ii] + TV[a * N3 + ii] * TV[a * N3 + ii]);
                    const int b = (a + 1) % 3, c = (a + 2) % 3;
                    const double curl_u = (DRV(U, b, c) - DRV(U, c, b)) * inv2dx;
                    const double curl_v = (DRV(V, b, c) - DRV(V, c, b)) * inv2dx;
                    Ex += -eta * (TU[a * N3 + ii] * curl_u + TV[a * N3 + ii] * curl_v);
                }
            }
            const double EQJ = (Eg + Em + Ep + Ex) * dV + 0.5 * (omega * Q - Omega * Jt);
            const double Jact = omega * ell - Omega * Kk;
            if (nrows < maxrows) {
                double *r = rows + 9 * nrows;
                r[0] = it; r[1] = omega; r[2] = Omega; r[3] = EQJ; r[4] = maxF; r[5] = smax; r[6] = Nn;
                r[7] = Jact; r[8] = cond;
            }
            nrows++;
        }
#undef LAP
#undef DRV
    }
#undef WR
#undef IX3
    double el, kk;                                            /* eta_qflow.c:377-399 */
    qflow_jhat(p, cur, jh, NULL, &el, &kk);
    ell = el * dV; Kk = kk * dV;
    if (cur != st) memcpy(st, cur, sizeof(double) * 12 * (size_t)N3);
    if (jhat_out) memcpy(jhat_out, jh, sizeof(double) * 12 * (size_t)N3);
    free(nx); free(jh); free(wmask);
    if (out) { out[0] = omega; out[1] = Omega; out[2] = cond; out[3] = ell; out[4] = Kk; }
    return nrows;
}
