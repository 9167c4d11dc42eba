/* scp_config.h — the sfa/sim configuration: every key the reference accepts
 * (sfa/sim/scp_config.h:217-343), its defaults (:165-215), its hard
 * validation errors (:439-542) and the KVMD key set embedded in snapshots
 * (sfa/sim/scp_init.h:505-556). Own table-driven implementation.
 *
 * Precedence (reference sfa/README.md:117-120, scp_sim.c:2716-2762):
 *   CLI -key value  >  cfg file  >  KVMD of a restart .sfa  >  defaults.
 */
#ifndef SCP_CONFIG_H
#define SCP_CONFIG_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int N; double L, T, dt_factor;
    double m2, mtheta2, eta, mu, kappa, kappa_h, alpha_cs, beta_h;
    int mode; double inv_alpha, inv_beta, kappa_gamma, gamma_conf, theta_sat, gamma_conv;
    double sigma_grad, chi_chiral, sigma_cubic, sigma_freq;
    double sigma_cross, lambda_self, theta_vev, tune_dt;
    int sweep; double sweep_T;
    int self_tune; double st_dt, st_eps, st_gamma, st_pcrit, st_charge; int st_project;
    double st_ptarget, st_gain;
    int complex_phi; char qball_profile[512]; double qball_omega, qball_x0, qball_y0, qball_z0;
    double qdiag_radius, qdiag_probe1, qdiag_probe2;
    int complex_gauge; double g_gauge, qball2_x0, qball2_y0, qball2_z0;
    int qball2_sign; double qball2_phase; int test_gauge_xform;
    double higgs_v, higgs_lam, higgs_kap, higgs_rvoid;
    double g_gauge2, gauge2_q0, gauge2_q1, gauge2_q2;
    int bc_type; double damp_width, damp_rate, bc_switch_time;
    double gradient_A_high, gradient_A_low; int gradient_margin;
    char init[32]; double A, sigma, A_bg, ellip, R_tube, delta[3];
    char init_sfa[512]; int init_frame; char init_exec[1024];
    char output[512]; char diag_file[512]; int precision; int output_split;
    double snap_dt, diag_dt, burst_start, burst_end, burst_every;
    double vec_snap_dt; int vec_iframe_interval; double vec_delta_tol; int vec_block_size;
    int n_fabrics, mf_stage, mf_lock_CQ; double eps_CQ, q_C, q_Q, q_L;
    char init_sfa_L[512], init_sfa_Q[512]; int mf_seed_Q;
    int n_locks; char locks_file[512], lock0[256], lock1[256], lock2[256], lock3[256];
    char locks_track[512]; double lock_soft_r, lock_soft_k, lock_bag_r, lock_bag_k;
    int lock_bag_mode, lock_bag_smooth, locks_medium_only;
    /* this implementation only: 0 = FP64 step, 1 = FP32 compensated step (key "step_precision") */
    int step_precision;
} scp_config;

void scp_cfg_defaults(scp_config *c);
/* Returns 0 when the key is known, 1 for an unknown key (a warning is printed,
 * as in the reference). */
int  scp_cfg_set(scp_config *c, const char *key, const char *val);
/* Returns 0, or -1 when the file cannot be opened. */
int  scp_cfg_load(scp_config *c, const char *path);
void scp_cfg_print(const scp_config *c);
/* Hard incompatibilities of the reference; writes "ERROR: ..." into err and
 * returns -1 when the config is refused. */
int  scp_cfg_validate(const scp_config *c, char *err, int errlen);
/* Sectors of the reference that this implementation does not run (multi-
 * fabric, Higgs, second gauge, locks, GPU-only real-mode experiments, vector
 * frames, self-tune...). Returns -1 with a
 * message naming the key when the config needs one of them. */
int  scp_cfg_unsupported(const scp_config *c, char *err, int errlen);
/* KVMD set 0 (25/27/29 keys, scp_init.h:505-556). Returns the pair count;
 * keys/vals point into `storage` (>= 64*32 chars). */
int  scp_cfg_kvmd(const scp_config *c, const char **keys, const char **vals, char *storage);
/* float -> half as the reference's f64_to_f16 (truncating, no subnormals,
 * scp_config.h:548-558) and back (:560-570). */
unsigned short scp_f64_to_f16(double v);
double scp_f16_to_f64(unsigned short h);

#ifdef __cplusplus
}
#endif
#endif
