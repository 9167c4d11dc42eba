/* scp_config.c — table-driven configuration (see scp_config.h). */
#include "scp_config.h"

#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

enum { K_INT, K_DBL, K_SQ, K_STR, K_PREC, K_DELTA, K_NOP };

typedef struct { const char *key; int kind; size_t off; size_t len; } keydef;

#define I(name)   { #name, K_INT, offsetof(scp_config, name), 0 }
#define D(name)   { #name, K_DBL, offsetof(scp_config, name), 0 }
#define S(name)   { #name, K_STR, offsetof(scp_config, name), sizeof(((scp_config *)0)->name) }

/* Every key of the reference cfg_set (scp_config.h:217-343). */
static const keydef KEYS[] = {
    I(N), D(L), D(T), D(dt_factor),
    { "m", K_SQ, offsetof(scp_config, m2), 0 }, { "m_theta", K_SQ, offsetof(scp_config, mtheta2), 0 },
    D(eta), D(mu), D(kappa), D(kappa_h), D(alpha_cs), D(beta_h),
    I(mode), D(inv_alpha), D(inv_beta), D(kappa_gamma), D(gamma_conf), D(theta_sat), D(gamma_conv),
    D(sigma_grad), D(chi_chiral), D(sigma_cubic), D(sigma_freq),
    D(sigma_cross), D(lambda_self), D(theta_vev), D(tune_dt),
    I(self_tune), D(st_dt), D(st_eps), D(st_gamma), D(st_pcrit), D(st_charge), I(st_project),
    D(st_ptarget), D(st_gain),
    I(complex_phi), S(qball_profile), D(qball_omega), D(qball_x0), D(qball_y0), D(qball_z0),
    D(qdiag_radius), D(qdiag_probe1), D(qdiag_probe2),
    I(complex_gauge), D(g_gauge),
    D(higgs_v), D(higgs_lam), D(higgs_kap), D(higgs_rvoid),
    D(g_gauge2), D(gauge2_q0), D(gauge2_q1), D(gauge2_q2),
    D(qball2_x0), D(qball2_y0), D(qball2_z0), I(qball2_sign), D(qball2_phase), I(test_gauge_xform),
    I(sweep), D(sweep_T),
    I(bc_type), D(damp_width), D(bc_switch_time), D(damp_rate),
    D(gradient_A_high), D(gradient_A_low), I(gradient_margin),
    S(init), D(A), D(sigma), D(A_bg), D(ellip), D(R_tube),
    { "delta", K_DELTA, offsetof(scp_config, delta), 0 },
    S(init_sfa), I(init_frame), S(init_exec), S(output), S(diag_file),
    D(snap_dt), D(diag_dt), I(output_split),
    D(vec_snap_dt), I(vec_iframe_interval), D(vec_delta_tol), I(vec_block_size),
    D(burst_start), D(burst_end), D(burst_every),
    { "precision", K_PREC, offsetof(scp_config, precision), 0 },
    { "vec_output", K_NOP, 0, 0 }, { "vec_kframe_interval", K_NOP, 0, 0 }, { "vec_n_fields", K_NOP, 0, 0 },
    I(n_fabrics), I(mf_stage), I(mf_lock_CQ), D(eps_CQ), D(q_C), D(q_Q), D(q_L),
    S(init_sfa_Q), I(mf_seed_Q), S(init_sfa_L),
    I(n_locks), S(locks_file), S(lock0), S(lock1), S(lock2), S(lock3), S(locks_track),
    D(lock_soft_r), D(lock_soft_k), D(lock_bag_r), D(lock_bag_k),
    I(lock_bag_mode), I(lock_bag_smooth), I(locks_medium_only),
    I(step_precision),
};
#undef I
#undef D
#undef S

void scp_cfg_defaults(scp_config *c) {                  /* scp_config.h:165-215 */
    memset(c, 0, sizeof *c);
    c->N = 128; c->L = 10.0; c->T = 200.0; c->dt_factor = 0.025;
    c->m2 = 2.25; c->mtheta2 = 0.0; c->eta = 0.5; c->mu = -41.345; c->kappa = 50.0;
    c->inv_alpha = 2.25; c->inv_beta = 5.0; c->kappa_gamma = 2.0;
    c->sweep_T = 50;
    c->st_dt = 2.0; c->st_eps = 0.04; c->st_gamma = 0.12; c->st_pcrit = 2.0; c->st_project = 1; c->st_gain = 0.10;
    c->qball_omega = 1.39; c->qball_x0 = c->qball_y0 = c->qball_z0 = 1e30;
    c->qdiag_radius = 8.0; c->qdiag_probe1 = 8.0; c->qdiag_probe2 = 16.0;
    c->g_gauge = 0.05;
    c->qball2_x0 = c->qball2_y0 = c->qball2_z0 = 1e30; c->qball2_sign = 1;
    c->higgs_lam = 1.0; c->higgs_kap = 2.0;
    c->gauge2_q0 = 1.0; c->gauge2_q1 = 1.0; c->gauge2_q2 = -2.0;
    c->damp_width = 3.0; c->damp_rate = 0.01;
    c->gradient_A_high = 0.15; c->gradient_A_low = 0.05; c->gradient_margin = 3;
    strcpy(c->init, "oscillon");
    c->A = 0.8; c->sigma = 3.0; c->A_bg = 0.1; c->ellip = 0.3325; c->R_tube = 3.0;
    c->delta[0] = 0.0; c->delta[1] = 3.0005; c->delta[2] = 4.4325;
    c->init_frame = -1;
    strcpy(c->output, "output.sfa"); strcpy(c->diag_file, "diag.tsv");
    c->precision = 1;
    c->snap_dt = 5.0; c->diag_dt = 2.0;
    c->vec_iframe_interval = 100; c->vec_delta_tol = 0.01; c->vec_block_size = 8;
    c->n_fabrics = 1; c->mf_lock_CQ = 1; c->q_Q = 1.0; c->q_L = -1.0;
}

int scp_cfg_set(scp_config *c, const char *key, const char *val) {
    for (size_t q = 0; q < sizeof KEYS / sizeof KEYS[0]; q++) {
        const keydef *k = &KEYS[q];
        if (strcmp(k->key, key)) continue;
        char *base = (char *)c + k->off;
        switch (k->kind) {
        case K_INT: *(int *)base = atoi(val); break;
        case K_DBL: *(double *)base = atof(val); break;
        case K_SQ: { double m = atof(val); *(double *)base = m * m; break; }
        case K_STR: strncpy(base, val, k->len - 1); break;
        case K_PREC:
            if (!strcmp(val, "f16")) c->precision = 0;
            else if (!strcmp(val, "f32")) c->precision = 1;
            else if (!strcmp(val, "f64")) c->precision = 2;
            break;
        case K_DELTA: sscanf(val, "%lf,%lf,%lf", &c->delta[0], &c->delta[1], &c->delta[2]); break;
        default: break;
        }
        return 0;
    }
    fprintf(stderr, "WARNING: unknown config key '%s'\n", key);
    return 1;
}

int scp_cfg_load(scp_config *c, const char *path) {     /* line grammar of scp_config.h:345-367 */
    FILE *fp = fopen(path, "r");
    if (!fp) return -1;
    char line[2048];
    while (fgets(line, sizeof line, fp)) {
        char *p = line;
        while (*p == ' ' || *p == '\t') p++;
        if (*p == '#' || *p == '\n' || *p == '\0') continue;
        char *eq = strchr(p, '=');
        if (!eq) continue;
        *eq = '\0';
        char *key = p, *val = eq + 1;
        for (char *e = eq - 1; e > key && (*e == ' ' || *e == '\t'); e--) *e = '\0';
        while (*val == ' ' || *val == '\t') val++;
        for (char *e = val + strlen(val) - 1; e > val && (*e == '\n' || *e == '\r' || *e == ' '); e--) *e = '\0';
        char *hash = strchr(val, '#');
        if (hash) { *hash = '\0'; for (char *e = hash - 1; e > val && *e == ' '; e--) *e = '\0'; }
        scp_cfg_set(c, key, val);
    }
    fclose(fp);
    return 0;
}

void scp_cfg_print(const scp_config *c) {
    static const char *PN[3] = {"f16", "f32", "f64"};
    printf("=== scp_sim (MI355X): %s step ===\n",
           c->complex_gauge ? "gauged U(1)" : (c->complex_phi ? "complex" : "real"));
    printf("Grid:    N=%d L=%.1f T=%.0f dt_factor=%.4f\n", c->N, c->L, c->T, c->dt_factor);
    printf("Physics: m²=%.4f m_θ²=%.4f η=%.3f μ=%.3f κ=%.1f κ_h=%.3f α=%.3f β=%.3f\n",
           c->m2, c->mtheta2, c->eta, c->mu, c->kappa, c->kappa_h, c->alpha_cs, c->beta_h);
    if (c->complex_phi) {
        printf("Complex: 12-field U(1) kernel (complex_phi=1)\n");
        printf("Q-diag:  radius=%.2f probes=(%.2f,%.2f)\n", c->qdiag_radius, c->qdiag_probe1, c->qdiag_probe2);
        if (c->complex_gauge) printf("Gauge:   U(1) compact links, g=%.4f (complex_gauge=1)\n", c->g_gauge);
    }
    printf("Mode:    %d", c->mode);
    if (c->bc_type == 0) printf("\nBC:      absorbing sphere (width=%.1f rate=%.4f)\n", c->damp_width, c->damp_rate);
    else if (c->bc_type == 1) printf("\nBC:      gradient pinned (A_high=%.3f A_low=%.3f margin=%d)\n",
                                     c->gradient_A_high, c->gradient_A_low, c->gradient_margin);
    else printf("\nBC:      periodic\n");
    printf("Init:    %s", c->init);
    if (!strcmp(c->init, "sfa")) printf(" (%s frame=%d)", c->init_sfa, c->init_frame);
    int pi = (c->precision >= 0 && c->precision <= 2) ? c->precision : 1;
    printf("\nOutput:  %s (%s, snap=%.1f diag=%.1f)\n", c->output, PN[pi], c->snap_dt, c->diag_dt);
    printf("Step:    %s\n\n", c->step_precision ? "FP32 compensated" : "FP64");
}

static int fail(char *err, int n, const char *msg) { snprintf(err, (size_t)n, "%s", msg); return -1; }

int scp_cfg_validate(const scp_config *c, char *err, int n) {   /* scp_config.h:439-542 */
    char msg[256];
    if (c->complex_phi != 0) {
        const char *bad = NULL;
        if (c->alpha_cs != 0) bad = "alpha_cs"; else if (c->beta_h != 0) bad = "beta_h";
        else if (c->kappa_h != 0) bad = "kappa_h"; else if (c->mode != 0) bad = "mode";
        else if (c->self_tune != 0) bad = "self_tune"; else if (c->sigma_grad != 0) bad = "sigma_grad";
        else if (c->sigma_cubic != 0) bad = "sigma_cubic"; else if (c->sigma_freq != 0) bad = "sigma_freq";
        else if (c->sigma_cross != 0) bad = "sigma_cross"; else if (c->lambda_self != 0) bad = "lambda_self";
        else if (c->theta_vev != 0) bad = "theta_vev"; else if (c->theta_sat != 0) bad = "theta_sat";
        else if (c->gamma_conf != 0) bad = "gamma_conf"; else if (c->gamma_conv != 0) bad = "gamma_conv";
        else if (c->sweep != 0) bad = "sweep"; else if (c->tune_dt != 0) bad = "tune_dt";
        else if (c->vec_snap_dt != 0) bad = "vec_snap_dt";
        if (bad) { snprintf(msg, sizeof msg, "ERROR: complex_phi=1 is incompatible with %s", bad); return fail(err, n, msg); }
        if (!strcmp(c->init, "braid")) return fail(err, n, "ERROR: braid init not defined for complex mode");
    }
    if (c->n_locks > 0) {
        if (c->complex_phi == 0 || c->complex_gauge == 0)
            return fail(err, n, "ERROR: n_locks>0 requires complex_phi=1 and complex_gauge=1");
        if (c->g_gauge == 0.0) return fail(err, n, "ERROR: n_locks>0 requires g_gauge != 0");
        if (c->n_locks > 4 && c->locks_file[0] == '\0') return fail(err, n, "ERROR: n_locks>4 requires locks_file=...");
        if (c->locks_medium_only && c->n_fabrics == 3)
            return fail(err, n, "ERROR: locks_medium_only incompatible with n_fabrics=3");
    }
    if (c->complex_gauge != 0) {
        if (c->complex_phi == 0) return fail(err, n, "ERROR: complex_gauge=1 requires complex_phi=1");
        if (c->bc_type == 1) return fail(err, n, "ERROR: complex_gauge=1 is incompatible with bc_type=1 (pinning A is gauge-dependent)");
        if (c->bc_switch_time != 0) return fail(err, n, "ERROR: complex_gauge=1 is incompatible with bc_switch_time != 0");
        if (c->g_gauge < 0) return fail(err, n, "ERROR: complex_gauge=1 requires g_gauge >= 0 (flip charges via the seed, not g)");
    }
    if (!strcmp(c->init, "qball")) {
        if (c->complex_phi == 0) return fail(err, n, "ERROR: init=qball requires complex_phi=1");
        if (c->qball_profile[0] == '\0') return fail(err, n, "ERROR: init=qball requires a non-empty qball_profile");
    }
    if (c->n_fabrics != 1 && c->n_fabrics != 3) {
        snprintf(msg, sizeof msg, "ERROR: n_fabrics must be 1 or 3 (got %d)", c->n_fabrics);
        return fail(err, n, msg);
    }
    if (c->n_fabrics == 3) {
        if (c->complex_phi == 0 || c->complex_gauge == 0)
            return fail(err, n, "ERROR: n_fabrics=3 requires complex_phi=1 and complex_gauge=1");
        if (c->higgs_v > 0.0) return fail(err, n, "ERROR: n_fabrics=3 does not yet support higgs_v (B1)");
        if (c->g_gauge2 != 0.0) return fail(err, n, "ERROR: n_fabrics=3 does not yet support g_gauge2 (B1)");
    }
    return 0;
}

int scp_cfg_unsupported(const scp_config *c, char *err, int n) {
    char msg[256];
    const char *bad = NULL;
    if (c->n_fabrics != 1) bad = "n_fabrics";
    else if (c->complex_gauge && c->higgs_v > 0) bad = "higgs_v";
    else if (c->complex_gauge && c->g_gauge2 != 0) bad = "g_gauge2";
    else if (c->n_locks > 0) bad = "n_locks";
    else if (c->sweep) bad = "sweep";
    else if (c->tune_dt != 0) bad = "tune_dt";
    else if (c->bc_type < 0 || c->bc_type > 2) bad = "bc_type";
    else if (c->output_split) bad = "output_split";
    if (bad) {
        snprintf(msg, sizeof msg, "ERROR: %s is not implemented by this MI355X build", bad);
        return fail(err, n, msg);
    }
    if (c->vec_snap_dt > 0) {       /* vector frames: the fit needs 4 samples per axis; frame % interval */
        const int bs = c->vec_block_size > 0 ? c->vec_block_size : 8;
        if (bs < 4 || bs > 16 || bs > c->N) {
            snprintf(msg, sizeof msg, "ERROR: vec_block_size=%d must lie in [4, min(16, N)] in this MI355X build", bs);
            return fail(err, n, msg);
        }
        if (c->vec_iframe_interval < 1) return fail(err, n, "ERROR: vec_iframe_interval must be >= 1");
    }
    return 0;
}

int scp_cfg_kvmd(const scp_config *c, const char **keys, const char **vals, char *st) {
    static const char *K[29] = {"N", "L", "T", "dt_factor", "m", "m_theta", "eta", "mu", "kappa",
                                "kappa_h", "alpha_cs", "beta_h", "bc_switch_time",
                                "mode", "inv_alpha", "inv_beta", "kappa_gamma",
                                "damp_width", "damp_rate", "precision", "delta",
                                "bc_type", "gradient_A_high", "gradient_A_low", "gradient_margin",
                                "complex_phi", "qball_omega", "complex_gauge", "g_gauge"};
    static const char *PN[3] = {"f16", "f32", "f64"};
    char *v[29];
    for (int q = 0; q < 29; q++) { v[q] = st + 64 * q; keys[q] = K[q]; vals[q] = v[q]; }
    /* value formats of scp_init.h:510-540 */
    snprintf(v[0], 32, "%d", c->N);           snprintf(v[1], 32, "%f", c->L);
    snprintf(v[2], 32, "%f", c->T);           snprintf(v[3], 32, "%f", c->dt_factor);
    snprintf(v[4], 32, "%f", sqrt(c->m2));    snprintf(v[5], 32, "%f", sqrt(c->mtheta2));
    snprintf(v[6], 32, "%f", c->eta);         snprintf(v[7], 32, "%f", c->mu);
    snprintf(v[8], 32, "%f", c->kappa);       snprintf(v[9], 32, "%.6f", c->kappa_h);
    snprintf(v[10], 32, "%.6f", c->alpha_cs); snprintf(v[11], 32, "%.6f", c->beta_h);
    snprintf(v[12], 32, "%.6f", c->bc_switch_time);
    snprintf(v[13], 32, "%d", c->mode);
    snprintf(v[14], 32, "%.6f", c->inv_alpha); snprintf(v[15], 32, "%.6f", c->inv_beta);
    snprintf(v[16], 32, "%.6f", c->kappa_gamma);
    snprintf(v[17], 32, "%.6f", c->damp_width); snprintf(v[18], 32, "%.6f", c->damp_rate);
    snprintf(v[19], 32, "%s", PN[c->precision >= 0 && c->precision <= 2 ? c->precision : 1]);
    snprintf(v[20], 64, "%.6f,%.6f,%.6f", c->delta[0], c->delta[1], c->delta[2]);
    snprintf(v[21], 32, "%d", c->bc_type);
    snprintf(v[22], 32, "%.6f", c->gradient_A_high); snprintf(v[23], 32, "%.6f", c->gradient_A_low);
    snprintf(v[24], 32, "%d", c->gradient_margin);
    snprintf(v[25], 32, "%d", c->complex_phi); snprintf(v[26], 32, "%.6f", c->qball_omega);
    snprintf(v[27], 32, "%d", c->complex_gauge); snprintf(v[28], 32, "%.9f", c->g_gauge);
    return c->complex_gauge ? 29 : (c->complex_phi ? 27 : 25);
}

unsigned short scp_f64_to_f16(double v) {              /* truncating, as scp_config.h:548-558 */
    float f = (float)v;
    uint32_t x; memcpy(&x, &f, 4);
    uint16_t sign = (x >> 16) & 0x8000;
    int e = (int)((x >> 23) & 0xFF) - 127 + 15;
    uint16_t mant = (x >> 13) & 0x3FF;
    if (e <= 0) return sign;
    if (e >= 31) return sign | 0x7C00;
    return (unsigned short)(sign | (e << 10) | mant);
}

double scp_f16_to_f64(unsigned short h) {
    uint16_t sign = h & 0x8000;
    int e = (h >> 10) & 0x1F;
    uint16_t mant = h & 0x3FF;
    if (e == 0) return sign ? -0.0 : 0.0;
    if (e == 31) return sign ? -INFINITY : INFINITY;
    uint32_t x = ((uint32_t)sign << 16) | ((uint32_t)(e - 15 + 127) << 23) | ((uint32_t)mant << 13);
    float f; memcpy(&f, &x, 4);
    return (double)f;
}
