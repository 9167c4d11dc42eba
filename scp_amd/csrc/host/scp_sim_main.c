/* scp_sim_main.c — the drop-in `scp_sim` binary on MI355X.
 *
 * Same command line, config keys, console lines, diag.tsv columns and SFA
 * snapshot layout as the reference CPU binary (sfa/sim/scp_sim.c:2706-3308);
 * the step, diagnostics and Gauss projection run on the GPU through the
 * C-ABI of include/scp_amd.h.
 *
 *   scp_sim config.cfg [-key value ...]
 *   scp_sim input.sfa  [-key value ...]     restart: KVMD config, N/L from the header, last frame
 *
 * Host-only work (init, SFA compression) uses OpenMP with OMP_NUM_THREADS
 * (default 4, as the reference).
 */
#define _GNU_SOURCE
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "scp_amd.h"
#include "scp_config.h"
#include "scp_init.h"
#include "sfa_io.h"
#include "vecframes.h"

static void die(const char *msg) {
    fprintf(stderr, "%s\n", msg);
    exit(1);
}

static void die_abi(const char *what) {
    char buf[512];
    scp_last_error(buf, sizeof buf);
    fprintf(stderr, "FATAL: %s: %s\n", what, buf);
    exit(1);
}

static const char *COLNAMES[30] = {
    "phi_x", "phi_y", "phi_z", "theta_x", "theta_y", "theta_z", "phi_vx", "phi_vy", "phi_vz",
    "theta_vx", "theta_vy", "theta_vz", "phiim_x", "phiim_y", "phiim_z", "thetaim_x", "thetaim_y",
    "thetaim_z", "phiim_vx", "phiim_vy", "phiim_vz", "thetaim_vx", "thetaim_vy", "thetaim_vz",
    "th_x", "th_y", "th_z", "E_x", "E_y", "E_z"};
static const int COLSEM[30] = {0, 0, 0, 1, 1, 1, 2, 2, 2, 2, 2, 2, 0, 0, 0, 1, 1, 1, 2, 2, 2, 2, 2, 2, 1, 1, 1, 2, 2, 2};
static const int COLCOMP[30] = {0, 1, 2, 0, 1, 2, 0, 1, 2, 3, 4, 5, 3, 4, 5, 3, 4, 5, 6, 7, 8, 9, 10, 11, 6, 7, 8, 12, 13, 14};

typedef struct {
    scp_grid *g;
    sfaio *out;
    double *state;       /* host state, snapshot column order */
    void *cast;          /* f32/f16 staging */
    int ncols, precision;
    long n3;
} Ctx;

static void download(Ctx *x) {
    double *cols[30];
    for (int c = 0; c < x->ncols; c++) cols[c] = x->state + (long)c * x->n3;
    if (scp_grid_download(x->g, cols, x->ncols)) die_abi("download");
}

/* sfa_snap (scp_sim.c:2646-2700) */
static void snap(Ctx *x, double t) {
    download(x);
    const void *cols[30];
    if (x->precision == 2) {
        for (int c = 0; c < x->ncols; c++) cols[c] = x->state + (long)c * x->n3;
    } else {
        int es = x->precision == 0 ? 2 : 4;
        for (int c = 0; c < x->ncols; c++) {
            const double *src = x->state + (long)c * x->n3;
            uint8_t *dst = (uint8_t *)x->cast + (size_t)c * x->n3 * es;
            if (x->precision == 1) {
                float *f = (float *)dst;
                #pragma omp parallel for schedule(static)
                for (long i = 0; i < x->n3; i++) f[i] = (float)src[i];
            } else {
                uint16_t *h = (uint16_t *)dst;
                #pragma omp parallel for schedule(static)
                for (long i = 0; i < x->n3; i++) h[i] = scp_f64_to_f16(src[i]);
            }
            cols[c] = dst;
        }
    }
    if (sfaio_write_frame(x->out, t, cols)) die("FATAL: snapshot write failed");
}

static double gauss_at_host(const scp_config *c, const double *s, long n3, long q, double dx) {
    const int N = c->N;
    int i = (int)(q / ((long)N * N)), j = (int)((q / N) % N), k = (int)(q % N);
    long nm[3] = {((long)((i - 1 + N) % N) * N + j) * N + k, ((long)i * N + (j - 1 + N) % N) * N + k,
                  ((long)i * N + j) * N + (k - 1 + N) % N};
    const double *E0 = s + 27 * n3, *E1 = s + 28 * n3, *E2 = s + 29 * n3;
    double divE = (E0[q] - E0[nm[0]] + E1[q] - E1[nm[1]] + E2[q] - E2[nm[2]]) / dx;
    double rho = 0;
    for (int a = 0; a < 3; a++)
        rho += s[a * n3 + q] * s[(18 + a) * n3 + q] - s[(12 + a) * n3 + q] * s[(6 + a) * n3 + q]
             + s[(3 + a) * n3 + q] * s[(21 + a) * n3 + q] - s[(15 + a) * n3 + q] * s[(9 + a) * n3 + q];
    return divE - c->g_gauge * rho;
}

/* do_test_gauge_xform (scp_sim.c:2461-2556), forces on the GPU */
static void test_gauge_xform(Ctx *x, const scp_config *c, double dx) {
    printf("\n=== test_gauge_xform: random lattice gauge transformation ===\n");
    const long n3 = x->n3;
    double *a0 = malloc(sizeof(double) * 15 * n3), *a1 = malloc(sizeof(double) * 15 * n3);
    double *G0 = malloc(sizeof(double) * n3);
    double *ac[15];
    for (int q = 0; q < 15; q++) ac[q] = a0 + (long)q * n3;
    if (scp_grid_forces(x->g, ac, 15)) die_abi("forces");
    download(x);
    for (long q = 0; q < n3; q++) G0[q] = gauss_at_host(c, x->state, n3, q, dx);
    scp_gauge_transform(c, x->state);
    const double *cols[30];
    for (int q = 0; q < 30; q++) cols[q] = x->state + (long)q * n3;
    if (scp_grid_upload(x->g, cols, 30)) die_abi("upload");
    for (int q = 0; q < 15; q++) ac[q] = a1 + (long)q * n3;
    if (scp_grid_forces(x->g, ac, 15)) die_abi("forces");
    const char *names[10] = {"|acc Phi_0|^2", "|acc Phi_1|^2", "|acc Phi_2|^2", "|acc Th_0|^2", "|acc Th_1|^2",
                             "|acc Th_2|^2", "E_acc_x", "E_acc_y", "E_acc_z", "G"};
    int pass = 1;
    for (int q = 0; q < 10; q++) {
        double sc = 0, dm = 0;
        for (long i = 0; i < n3; i++) {
            double b, a;
            if (q < 3) {
                b = a0[q * n3 + i] * a0[q * n3 + i] + a0[(6 + q) * n3 + i] * a0[(6 + q) * n3 + i];
                a = a1[q * n3 + i] * a1[q * n3 + i] + a1[(6 + q) * n3 + i] * a1[(6 + q) * n3 + i];
            } else if (q < 6) {
                int r = q - 3;
                b = a0[(3 + r) * n3 + i] * a0[(3 + r) * n3 + i] + a0[(9 + r) * n3 + i] * a0[(9 + r) * n3 + i];
                a = a1[(3 + r) * n3 + i] * a1[(3 + r) * n3 + i] + a1[(9 + r) * n3 + i] * a1[(9 + r) * n3 + i];
            } else if (q < 9) {
                b = a0[(12 + q - 6) * n3 + i];
                a = a1[(12 + q - 6) * n3 + i];
            } else {
                b = G0[i];
                a = gauss_at_host(c, x->state, n3, i, dx);
            }
            if (fabs(b) > sc) sc = fabs(b);
            if (fabs(a - b) > dm) dm = fabs(a - b);
        }
        double rel = dm / fmax(sc, 1.0);
        printf("  %-14s max|diff|=%.3e scale=%.3e rel=%.3e %s\n", names[q], dm, sc, rel, rel <= 1e-12 ? "ok" : "FAIL");
        if (rel > 1e-12) pass = 0;
    }
    printf("test_gauge_xform: %s\n", pass ? "PASS" : "FAIL");
    exit(pass ? 0 : 1);
}

int main(int argc, char **argv) {
    scp_config c;
    scp_cfg_defaults(&c);
    if (argc < 2) {
        fprintf(stderr, "Usage: %s config.cfg [-key value ...]\n", argv[0]);
        fprintf(stderr, "       %s input.sfa [-key value ...]   (restart from SFA)\n", argv[0]);
        return 1;
    }
    const char *a1 = argv[1];
    size_t l1 = strlen(a1);
    if (l1 > 4 && !strcmp(a1 + l1 - 4, ".sfa")) {          /* scp_sim.c:2720-2751 */
        printf("Loading parameters from SFA: %s\n", a1);
        sfaio *s = sfaio_open(a1);
        if (!s) { fprintf(stderr, "Cannot open SFA: %s\n", a1); return 1; }
        int nk = sfaio_kvmd_count(s);
        if (nk == 0) printf("  WARNING: no KVMD metadata found — using defaults\n");
        else {
            int use = 0;
            for (int i = 0; i < nk; i++) {
                uint32_t ff;
                const char *k, *v;
                sfaio_kvmd_pair(s, i, -1, &k, &v, &ff);
                if (ff == 0xFFFFFFFFu) { use = i; break; }
            }
            const char *k, *v;
            uint32_t ff;
            int np = sfaio_kvmd_pair(s, use, -1, &k, &v, &ff);
            printf("  KVMD set %d: %d parameters\n", (int)s->kvmd[use].set_id, np);
            for (int p = 0; p < np; p++) {
                sfaio_kvmd_pair(s, use, p, &k, &v, &ff);
                scp_cfg_set(&c, k, v);
            }
        }
        char tmp[64];
        snprintf(tmp, 64, "%u", s->Nx); scp_cfg_set(&c, "N", tmp);
        snprintf(tmp, 64, "%.6f", s->Lx); scp_cfg_set(&c, "L", tmp);
        scp_cfg_set(&c, "init", "sfa");
        strncpy(c.init_sfa, a1, 511);
        c.init_frame = -1;
        sfaio_close(s);
    } else if (scp_cfg_load(&c, a1)) {
        fprintf(stderr, "Cannot open config: %s\n", a1);
        exit(1);
    }
    for (int i = 2; i < argc - 1; i += 2) {                 /* CLI overrides, scp_sim.c:2758-2762 */
        const char *key = argv[i];
        if (key[0] == '-') key++;
        scp_cfg_set(&c, key, argv[i + 1]);
    }
    int nth = 4;
    const char *env = getenv("OMP_NUM_THREADS");
    if (env) nth = atoi(env);
    if (nth < 1) nth = 1;
    omp_set_num_threads(nth);

    scp_cfg_print(&c);
    char err[512];
    if (scp_cfg_validate(&c, err, sizeof err)) die(err);
    if (scp_cfg_unsupported(&c, err, sizeof err)) die(err);

    const int mode = c.complex_gauge ? 2 : (c.complex_phi ? 1 : 0);
    const int ncols = mode == 2 ? 30 : (mode == 1 ? 24 : 12);
    const double dx = 2.0 * c.L / (c.N - 1), dt = c.dt_factor * dx;   /* scp_sim.c:204-205 */
    const long n3 = (long)c.N * c.N * c.N;
    scp_params p;
    memset(&p, 0, sizeof p);
    p.N = c.N; p.mode = mode; p.precision = c.step_precision ? SCP_PREC_F32C : SCP_PREC_F64;
    p.lo_words = 4;   /* FP32C: int8 compensation words on the positions (the production form) */
    p.bc_type = c.bc_type; p.rmode = c.mode;
    p.gradient_margin = c.bc_type == 1 ? c.gradient_margin : 0;
    p.L = c.L; p.dx = dx; p.dt = dt;
    p.m2 = c.m2; p.mtheta2 = c.mtheta2; p.eta = c.eta; p.mu = c.mu; p.kappa = c.kappa; p.g = c.g_gauge;
    p.kappa_h = c.kappa_h; p.alpha_cs = c.alpha_cs; p.beta_h = c.beta_h;
    p.inv_alpha = c.inv_alpha; p.inv_beta = c.inv_beta; p.kappa_gamma = c.kappa_gamma;
    p.damp_width = c.damp_width; p.damp_rate = c.damp_rate;
    p.qdiag_radius = c.qdiag_radius; p.qdiag_probe1 = c.qdiag_probe1; p.qdiag_probe2 = c.qdiag_probe2;
    /* real-mode terms of the reference's GPU step (scp_sim.cu:2930-2937); theta_vev and
       gamma_conf are parsed and never used by the reference, so they change nothing here */
    p.theta_sat = c.theta_sat; p.gamma_conv = c.gamma_conv; p.sigma_grad = c.sigma_grad;
    p.chi_chiral = c.chi_chiral; p.sigma_cubic = c.sigma_cubic; p.sigma_freq = c.sigma_freq;
    p.sigma_cross = c.sigma_cross; p.lambda_self = c.lambda_self;
    int64_t hbm = scp_grid_bytes(&p);
    if (hbm < 0) die_abi("parameters");
    printf("Allocating %.2f GB host + %.2f GB HBM (N=%d, %d fields%s)\n", ncols * n3 * 8.0 / 1e9, hbm / 1e9,
           c.N, mode ? 12 : 6, mode == 2 ? " + U(1) links" : "");
    printf("dx=%.4f dt=%.6f threads=%d\n\n", dx, dt, nth);

    Ctx x;
    memset(&x, 0, sizeof x);
    x.ncols = ncols; x.n3 = n3; x.precision = c.precision;
    x.state = malloc(sizeof(double) * (size_t)ncols * (size_t)n3);
    if (!x.state) die("FATAL: malloc failed");
    if (scp_do_init(&c, mode, x.state, err, sizeof err)) die(err);

    const int gauged = (c.complex_gauge && c.g_gauge != 0.0);
    if (gauged && c.bc_type == 2) {                          /* scp_sim.c:2264-2279 */
        double qn, qa;
        scp_net_charge(&c, mode, x.state, &qn, &qa);
        if (fabs(qn) > 1e-6 * fmax(qa, 1.0)) {
            fprintf(stderr, "ERROR: bc_type=2 (periodic) requires net-neutral seed "
                    "(Q_net=%.3e, Q_abs=%.3e); use +/- pairs or bc_type=0. "
                    "jellium=1 is deferred (not implemented in v1).\n", qn, qa);
            exit(1);
        }
    }
    if (scp_grid_create(&p, &x.g)) die_abi("grid create");
    {
        const double *cols[30];
        for (int q = 0; q < ncols; q++) cols[q] = x.state + (long)q * n3;
        if (scp_grid_upload(x.g, cols, ncols)) die_abi("upload");
    }
    if (gauged) {
        int it; double gres;
        if (scp_grid_gauss_project(x.g, &it, &gres)) die_abi("Gauss projection");
        printf("Gauss projection: %d CG iters, gauss_max(0)=%.2e (offset %.3e)\n", it, gres, scp_grid_g_offset(x.g));
    }
    if (c.test_gauge_xform) {
        if (!gauged) die("ERROR: test_gauge_xform=1 requires complex_gauge=1 and g_gauge != 0");
        test_gauge_xform(&x, &c, dx);
    }
    if (c.bc_type == 1)   /* the uploaded state is the pinned state (scp_sim.c:2845-2849) */
        printf("Gradient BC: pinned %d slabs on each x-face (A_high=%.3f, A_low=%.3f)\n\n",
               c.gradient_margin, c.gradient_A_high, c.gradient_A_low);

    int sfa_dtype = c.precision == 0 ? SFAIO_F16 : (c.precision == 1 ? SFAIO_F32 : SFAIO_F64);
    x.out = sfaio_create(c.output, c.N, c.N, c.N, c.L, c.L, c.L, dt);
    if (!x.out) {
        fprintf(stderr, "ERROR: cannot create output file '%s' (missing directory or unwritable path?)\n", c.output);
        exit(1);
    }
    int zst = sfaio_zstd_available();
    if (sfaio_set_codec(x.out, zst ? SFAIO_CODEC_COLZSTD : SFAIO_CODEC_RAW, 1))
        die("FATAL: cannot select the snapshot codec");
    if (!zst) printf("WARNING: libzstd not found; snapshots are written RAW\n");
    {
        const char *keys[32], *vals[32];
        char store[64 * 32];
        int nk = scp_cfg_kvmd(&c, keys, vals, store);
        sfaio_add_kvmd(x.out, 0, 0xFFFFFFFFu, 0xFFFFFFFFu, keys, vals, nk);
    }
    for (int q = 0; q < ncols; q++) sfaio_add_column(x.out, COLNAMES[q], sfa_dtype, COLSEM[q], COLCOMP[q]);
    sfaio_finalize_header(x.out);
    static const char *PN[3] = {"f16", "f32", "f64"};
    printf("SFA: %s (%d cols, %s, %s)\n\n", c.output, ncols, PN[c.precision], zst ? "BSS+zstd" : "raw");
    if (c.precision != 2) x.cast = malloc((size_t)ncols * n3 * (c.precision == 0 ? 2 : 4));
    snap(&x, 0.0);

    /* vector frames (FRVD) into the same file, real mode (scp_sim.c:2942-2988) */
    scp_vecframes vf;
    int vec_every = 0;
    const int vec_on = c.vec_snap_dt > 0;
    if (vec_on) {
        if (!zst) die("ERROR: vector frames need libzstd (FRVD payloads are zstd-compressed)");
        vec_every = (int)lround(c.vec_snap_dt / dt);
        if (vec_every < 1) vec_every = 1;
        if (scp_vf_init(&vf, c.N, c.vec_block_size > 0 ? c.vec_block_size : 8, c.vec_iframe_interval, c.vec_delta_tol))
            die("FATAL: malloc failed (vector frames)");
        printf("Vec frames: BS=%d, %d patches, %d coeffs/patch, snap every %d steps (temporal)\n", vf.bs,
               vf.n_patches, vf.nc, vec_every);
    }

    FILE *fp = fopen(c.diag_file, "w");
    if (!fp) die("ERROR: cannot open diag file");
    fprintf(fp, "t\tE_phi_kin\tE_theta_kin\tE_grad\tE_mass\tE_pot\tE_tgrad\tE_tmass\t"
                "E_coupling\tE_total\tphi_max\tP_max\tP_int\ttheta_rms");
    if (c.complex_phi) fprintf(fp, "\tQ_phi\tQ_theta\tQ_total\ts_max\tr_core\tomega_core\tQ_core\tthp1u\tthp1v\tthp2u\tthp2v");
    if (c.complex_gauge) fprintf(fp, "\tgauss_max\tgauss_l2\tE_em\tQ_flux");
    if (c.complex_phi) fprintf(fp, "\tQ_p0\tQ_p1\tQ_p2");
    fprintf(fp, "\n");

    const int n_steps = (int)lround(c.T / dt);
    int diag_every = (int)lround(c.diag_dt / dt); if (diag_every < 1) diag_every = 1;
    int snap_every = (int)lround(c.snap_dt / dt); if (snap_every < 1) snap_every = 1;
    const double burst_dur = c.burst_end - c.burst_start;
    if (c.burst_start > 0 && burst_dur > 0) {
        if (c.burst_every > 0)
            printf("Burst snap: %.1f duration every %.0f time units starting t=%.1f\n", burst_dur, c.burst_every, c.burst_start);
        else
            printf("Burst snap: every timestep from t=%.1f to t=%.1f\n", c.burst_start, c.burst_end);
    }

    double d[SCP_D_COUNT];
    if (scp_grid_diag(x.g, d, SCP_D_COUNT)) die_abi("diag");
    const double E0 = d[SCP_D_E_TOTAL], trms0 = d[SCP_D_THETA_RMS];
    printf("INIT: E_total=%.4e E_pot=%.4f phi_max=%.4f P_int=%.4e theta_rms=%.3e\n\n",
           d[SCP_D_E_TOTAL], d[SCP_D_E_POT], d[SCP_D_PHI_MAX], d[SCP_D_P_INT], d[SCP_D_THETA_RMS]);
#define ROW(t)                                                                                          \
    do {                                                                                                \
        fprintf(fp, "%.2f\t%.6e\t%.6e\t%.6e\t%.6e\t%.6e\t%.6e\t%.6e\t%.6e\t%.6e\t%.6e\t%.6e\t%.6e\t%.6e", \
                (t), d[0], d[1], d[2], d[3], d[4], d[5], d[6], d[7], d[8], d[9], d[10], d[11], d[12]);  \
        if (c.complex_phi)                                                                              \
            fprintf(fp, "\t%.12e\t%.12e\t%.12e\t%.6e\t%.6e\t%.12e\t%.12e\t%.6e\t%.6e\t%.6e\t%.6e",       \
                    d[SCP_D_Q_PHI], d[SCP_D_Q_THETA], d[SCP_D_Q_PHI] + d[SCP_D_Q_THETA], d[SCP_D_S_MAX],  \
                    d[SCP_D_R_CORE], d[SCP_D_OMEGA_CORE], d[SCP_D_Q_CORE], d[SCP_D_THP1U],               \
                    d[SCP_D_THP1V], d[SCP_D_THP2U], d[SCP_D_THP2V]);                                    \
        if (c.complex_gauge)                                                                            \
            fprintf(fp, "\t%.6e\t%.6e\t%.12e\t%.12e", d[SCP_D_GAUSS_MAX], d[SCP_D_GAUSS_L2],            \
                    d[SCP_D_E_EM], d[SCP_D_Q_FLUX]);                                                    \
        if (c.complex_phi) fprintf(fp, "\t%.12e\t%.12e\t%.12e", d[SCP_D_Q_P0], d[SCP_D_Q_P1], d[SCP_D_Q_P2]); \
        fprintf(fp, "\n");                                                                              \
    } while (0)
    ROW(0.0);

    /* self-tune (the reference's GPU binary, scp_sim.cu:5475-5547; real mode): kappa
       follows the stability edge at fixed charge, phi rescaled to the target charge */
    const int st_every = (c.self_tune && c.st_dt > 0) ? (int)lround(c.st_dt / dt) : 0;
    double st_Q0 = -1.0;
    if (st_every > 0)
        printf("Self-tune: kappa dynamical, every %d steps (%.1f t.u.); "
               "eps=%.3f gamma=%.3f pcrit=%.2f project=%d  kappa0=%.3f\n\n",
               st_every, c.st_dt, c.st_eps, c.st_gamma, c.st_pcrit, c.st_project, c.kappa);

    double wall0 = omp_get_wtime();
    int major = diag_every * 25; if (major < 1) major = 1;
    if (major > n_steps / 10) { major = (n_steps / 10 / diag_every) * diag_every; if (major < diag_every) major = diag_every; }
    int pending = 0;
    /* absorbing -> periodic at bc_switch_time: damp_rate = 0 from that step on
       (scp_sim.c:3073-3081, scp_sim.cu:5464-5490) */
    const int bc_switch_step = (c.bc_switch_time > 0) ? (int)lround(c.bc_switch_time / dt) : 0;
    int bc_switched = 0;
    for (int step = 1; step <= n_steps; step++) {
        double t = step * dt;
        if (bc_switch_step > 0 && step == bc_switch_step && !bc_switched) {
            if (pending && scp_grid_step(x.g, pending)) die_abi("step");   /* steps before the switch keep the sponge */
            pending = 0;
            if (scp_grid_set_damp_rate(x.g, 0.0)) die_abi("damp rate");
            c.damp_rate = 0.0;
            printf("\n*** BC SWITCH at t=%.1f: absorbing -> periodic ***\n\n", t);
            bc_switched = 1;
        }
        int in_burst = 0;
        if (burst_dur > 0) {
            double tr = t - c.burst_start;
            if (tr >= 0) {
                if (c.burst_every > 0) tr = fmod(tr, c.burst_every);
                in_burst = (tr >= 0 && tr < burst_dur);
            }
        }
        const int do_snap = in_burst || step % snap_every == 0;
        const int do_diag = step % diag_every == 0;
        const int do_vec = vec_on && step % vec_every == 0;
        const int do_st = st_every > 0 && step % st_every == 0;
        pending++;
        if (!do_snap && !do_diag && !do_vec && !do_st && step != n_steps) continue;
        /* a diag step takes the diagnostics fused into its last launch (unless a
           self-tune rescale comes between the step and the diag row) */
        if (do_diag && !do_st) {
            if (scp_grid_step_diag(x.g, pending, d, SCP_D_COUNT)) die_abi("step");
        } else if (scp_grid_step(x.g, pending)) {
            die_abi("step");
        }
        pending = 0;
        if (do_st) {                                          /* scp_sim.cu:5502-5547 */
            double sd[SCP_D_COUNT];
            if (scp_grid_diag(x.g, sd, SCP_D_COUNT)) die_abi("diag");
            const double Pmax = sd[SCP_D_P_MAX], Emass = sd[SCP_D_E_MASS];
            const double Q = (c.m2 > 1e-12) ? 2.0 * Emass / c.m2 : Emass;
            if (st_Q0 < 0) st_Q0 = (c.st_charge > 0 ? c.st_charge : Q);   /* the target, set once */
            if (c.st_project && c.st_ptarget <= 0 && Q > 1e-12)             /* charge projection */
                if (scp_grid_rescale_phi(x.g, sqrt(st_Q0 / Q))) die_abi("rescale");
            const int nanflag = (Pmax != Pmax) || (Pmax > 1e30);
            if (c.st_ptarget > 0.0) {                                       /* density homeostasis */
                double err = nanflag ? 1.0 : (Pmax - c.st_ptarget) / c.st_ptarget;
                if (err > 1.0) err = 1.0;
                if (err < -1.0) err = -1.0;
                c.kappa *= (1.0 + c.st_gain * err);
                if (c.kappa < 1e-3) c.kappa = 1e-3;
                if (scp_grid_set_kappa(x.g, c.kappa)) die_abi("kappa");
                printf("[self-tune] t=%6.1f kappa=%9.4f Pmax=%10.3e target=%.4f err=%+.3f Q=%9.3e\n",
                       t, c.kappa, Pmax, c.st_ptarget, err, Q);
            } else {                                                         /* SOC update */
                const int collapse = nanflag || (Pmax > c.st_pcrit);
                c.kappa *= collapse ? (1.0 + c.st_gamma) : (1.0 - c.st_eps);
                if (c.kappa < 1e-3) c.kappa = 1e-3;
                if (scp_grid_set_kappa(x.g, c.kappa)) die_abi("kappa");
                printf("[self-tune] t=%6.1f kappa=%9.4f Pmax=%10.3e Q=%9.3e %s\n", t, c.kappa, Pmax, Q,
                       collapse ? "COLLAPSE" : "stable");
            }
            fflush(stdout);
            if (do_diag && scp_grid_diag(x.g, d, SCP_D_COUNT)) die_abi("diag");
        }
        if (do_snap) snap(&x, t);
        if (do_vec) {                                         /* scp_sim.c:3098-3193 */
            if (scp_grid_vec_fit(x.g, vf.bs, vf.coeffs, (int64_t)vf.n_patches * vf.nc)) die_abi("vector fit");
            if (scp_vf_write(&vf, x.out, t)) die("FATAL: cannot write a vector frame");
        }
        if (do_diag) {
            ROW(t);
            fflush(fp);
            if (step % major == 0) {
                double wall = omp_get_wtime() - wall0, et = d[SCP_D_E_TOTAL];
                printf("t=%7.1f E=%.3e (drift %+.3f%%) Ep=%.1f phi=%.3f θ_rms=%.2e ", t, et,
                       100 * (et - E0) / (fabs(E0) + 1e-30), d[SCP_D_E_POT], d[SCP_D_PHI_MAX], d[SCP_D_THETA_RMS]);
                if (c.complex_phi) printf("Q=%.6e s_max=%.4e ", d[SCP_D_Q_PHI] + d[SCP_D_Q_THETA], d[SCP_D_S_MAX]);
                if (c.complex_gauge) printf("gauss=%.1e E_em=%.3e ", d[SCP_D_GAUSS_MAX], d[SCP_D_E_EM]);
                printf("[%.0f%% %.1fs %.2fms/step]\n", 100.0 * step / n_steps, wall, 1000 * wall / step);
                fflush(stdout);
            }
        }
    }
    {                                                         /* final frame rule, scp_sim.c:3254-3259 */
        int last = (n_steps / snap_every) * snap_every, gap = n_steps - last;
        if (gap > snap_every / 2) snap(&x, n_steps * dt);
    }
    uint32_t nf = x.out->total_frames;
    sfaio_close(x.out);
    if (vec_on) {
        printf("Vec frames: %d written to %s (temporal)\n", vf.frame, c.output);
        scp_vf_free(&vf);
    }
    if (scp_grid_diag(x.g, d, SCP_D_COUNT)) die_abi("diag");
    double wall = omp_get_wtime() - wall0;
    double et = d[SCP_D_E_TOTAL], Qp = d[SCP_D_Q_PHI], Qt = d[SCP_D_Q_THETA];
    printf("\n=== COMPLETE ===\n");
    printf("E_total=%.4e (drift %.3f%%) E_pot=%.4f\n", et, 100 * (et - E0) / (fabs(E0) + 1e-30), d[SCP_D_E_POT]);
    printf("phi_max=%.4f P_int=%.4e theta_rms=%.3e\n", d[SCP_D_PHI_MAX], d[SCP_D_P_INT], d[SCP_D_THETA_RMS]);
    if (c.complex_phi) {
        printf("Q_phi=%.6e Q_theta=%.6e Q_total=%.6e s_max=%.4e r_core=%.3f omega_core=%.4f Q_core=%.6e\n",
               Qp, Qt, Qp + Qt, d[SCP_D_S_MAX], d[SCP_D_R_CORE], d[SCP_D_OMEGA_CORE], d[SCP_D_Q_CORE]);
        printf("Q_p = (%.6e, %.6e, %.6e)\n", d[SCP_D_Q_P0], d[SCP_D_Q_P1], d[SCP_D_Q_P2]);
    }
    if (c.complex_gauge)
        printf("gauss_max=%.3e gauss_l2=%.3e E_em=%.6e Q_flux=%.6e\n", d[SCP_D_GAUSS_MAX], d[SCP_D_GAUSS_L2],
               d[SCP_D_E_EM], d[SCP_D_Q_FLUX]);
    printf("SFA: %s (%u frames)\n", c.output, nf);
    double trms = d[SCP_D_THETA_RMS];
    printf("[%s] theta_rms grew: %.2e -> %.2e\n", (trms > trms0 + 1e-10) ? "OK" : "WARN", trms0, trms);
    printf("Wall: %.1fs (%.1f min) %.2fms/step\n", wall, wall / 60, 1000 * wall / (n_steps > 0 ? n_steps : 1));
    fclose(fp);
    scp_grid_destroy(x.g);
    free(x.state);
    free(x.cast);
    return 0;
}
