/* scp_oracle.h — CPU FP64 restatement of the sfa/sim leapfrog step.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: tests/, smoke() and
 * the cpu_baseline leg of bench.py may load it; the product (scp_amd) never
 * links, loads or calls it.
 *
 * State layout (shared with the product's C-ABI): one contiguous buffer of
 * ncol*N^3 doubles in the SFA snapshot column order of the reference
 * (sfa/sim/scp_sim.c:2862-2897):
 *   0-2  phi_x..z        (u_a, Re Phi_a)
 *   3-5  theta_x..z      (tu_a)
 *   6-8  phi_vx..vz      (udot_a)
 *   9-11 theta_vx..vz    (tudot_a)
 *  12-14 phiim_x..z      (v_a)            complex/gauged only
 *  15-17 thetaim_x..z    (tv_a)
 *  18-20 phiim_vx..vz    (vdot_a)
 *  21-23 thetaim_vx..vz  (tvdot_a)
 *  24-26 th_x..z         (link angles)    gauged only
 *  27-29 E_x..z          (link E fields)
 * Site index idx = i*N*N + j*N + k (k fastest), reference scp_sim.c:758.
 *
 * Acceleration buffer: nacc*N^3 doubles:
 *   0-2 phi_acc, 3-5 theta_acc, [6-8 phi_im_acc, 9-11 theta_im_acc], [12-14 E_acc]
 */
#ifndef SCP_ORACLE_H
#define SCP_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int    N;
    int    mode;          /* 0 real, 1 complex, 2 gauged */
    int    bc_type;       /* 0 absorbing sphere, 1 gradient pinned, 2 periodic */
    int    rmode;         /* real-mode mass coupling: 0 const, 1 inverse, 3 density-kappa */
    double L, dx, dt;
    double m2, mtheta2, eta, mu, kappa, g;
    double kappa_h, alpha_cs, beta_h;
    double inv_alpha, inv_beta, kappa_gamma;
    double damp_width, damp_rate;
    double G_offset;      /* frozen jellium offset from the Gauss projection */
    double qdiag_radius, qdiag_probe1, qdiag_probe2;
    int    gradient_margin;   /* bc 1: pinned x-slabs per face */
    const double *pin;        /* bc 1: the pinned state (same layout as the state) */
    /* real mode: terms of the reference's GPU step only (scp_sim.cu:1144-1248) */
    double theta_sat, gamma_conv, sigma_grad, chi_chiral, sigma_cubic, sigma_freq, sigma_cross, lambda_self;
} OrParams;

/* Diagnostics vector: every column of the reference diag.tsv row for one time
 * (scp_sim.c:3040-3065), in this fixed order. */
enum {
    ORD_E_PHI_KIN = 0, ORD_E_THETA_KIN, ORD_E_GRAD, ORD_E_MASS, ORD_E_POT,
    ORD_E_TGRAD, ORD_E_TMASS, ORD_E_COUPLING, ORD_E_TOTAL, ORD_PHI_MAX,
    ORD_P_MAX, ORD_P_INT, ORD_THETA_RMS,
    ORD_Q_PHI, ORD_Q_THETA, ORD_Q_TOTAL, ORD_S_MAX, ORD_R_CORE, ORD_OMEGA_CORE,
    ORD_Q_CORE, ORD_THP1U, ORD_THP1V, ORD_THP2U, ORD_THP2V,
    ORD_GAUSS_MAX, ORD_GAUSS_L2, ORD_E_EM, ORD_Q_FLUX,
    ORD_Q_P0, ORD_Q_P1, ORD_Q_P2,
    ORD_COUNT
};

int    or_ncols(int mode);
int    or_nacc(int mode);
void   or_forces(const OrParams *p, const double *state, double *acc);
void   or_step(const OrParams *p, double *state, double *acc, int nsteps);
void   or_diag(const OrParams *p, const double *state, double *out);
double or_net_charge(const OrParams *p, const double *state, double *q_abs);
int    or_gauss_project(OrParams *p, double *state, double *gauss_max0);
void   or_gauge_transform(const OrParams *p, double *state);

/* Fixed-charge gradient flow (sfa/seed/eta_qflow.c, plain fixed-Q mode).
 * state: 12 N^3 doubles u0..2, v0..2, tu0..2, tv0..2 (site i*N*N + j*N + k). */
typedef struct {
    int    N;
    double L, eta, m, mth, mu, kappa;
} OrQflow;
/* N = int sum_a (|phi_a|^2 + |theta_a|^2), summed in site order (eta_qflow.c:193-197) */
double or_qflow_norm(const OrQflow *p, const double *state);
/* Iterations it = 0..iters (eta_qflow.c:242-375): omega = Q/N, Jacobi step
 * x += dtau F. A row {it, omega, E_Q, maxF, s_max, N} for it % printev == 0 and
 * it == iters, at most maxrows; returns the row count; *omega_last = the
 * multiplier of the last iteration (the one the output velocities use). */
int    or_qflow(const OrQflow *p, double *state, double Q, double dtau, int iters, int printev, double *rows,
                int maxrows, double *omega_last);
/* The fixed-(Q, J) mode (eta_qflow.c fixJ = 1): Jhat w = x d_y w - y d_x w - zhat x w
 * of the 12 fields (:210-262), the two multipliers (omega, Omega) from the
 * regularised 2x2 solve with |Omega| capped at 0.15, the Coriolis and centrifugal
 * forces masked at the light cylinder (:294-310). A row {it, omega, Omega, E_QJ,
 * maxF, s_max, N, J_act, cond} for it % printev == 0 and it == iters (:352-358);
 * returns the row count. out[0..5): omega, Omega, cond of the last iteration,
 * ell and K of the final state (the refresh :377-399); jhat (12 N^3, or NULL):
 * Jhat of the final state's fields. */
int    or_qflow_j(const OrQflow *p, double *state, double Q, double Jt, double dtau, int iters, int printev,
                  double *rows, int maxrows, double *out, double *jhat);
/* The light-cylinder mask of the rotational terms at column (i, j) (eta_qflow.c:213-217) */
double or_qflow_wmask(const OrQflow *p, int i, int j);

#ifdef __cplusplus
}
#endif
#endif
