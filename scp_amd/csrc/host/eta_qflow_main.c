/* eta_qflow_main.c — the drop-in fixed-charge flow relaxer (scp_amd/bin/eta_qflow).
 *
 * Same command line, printed lines and output file as sfa/seed/eta_qflow.c
 * (usage eta_qflow.c:72-86, header lines :199-206, rows :370, final lines
 * :400-432, the f32 rotating seed :436-468), with the flow itself on the GPU
 * through the C-ABI (scp_qflow_*, include/scp_amd.h) and the seed built by
 * libscp_io (qflow_seed.h): the ball and the ring/twisted seeds (wind, wtor,
 * twist), in the fixed-Q mode and the fixed-(Q, J) mode (fixJ = 1, Jtarget;
 * header :193-197, rows :353-357, final line :398-402, the seed's rotational
 * velocity terms :447-460).
 */
#include "qflow_seed.h"
#include "scp_amd.h"
#include "sfa_io.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static void die_abi(const char *what) {
    char buf[512];
    scp_last_error(buf, sizeof buf);
    fprintf(stderr, "FATAL: %s: %s\n", what, buf);
    exit(1);
}

int main(int argc, char **argv) {
    if (argc < 11) {
        fprintf(stderr, "usage: %s N L profile omega0 eta m mth mu kappa out.sfa [Q iters dtau thinit printev]\n",
                argv[0]);
        return 1;
    }
    const int N = atoi(argv[1]);
    const double L = atof(argv[2]);
    const char *prof = argv[3];
    const double omega0 = atof(argv[4]), eta = atof(argv[5]), m = atof(argv[6]), mth = atof(argv[7]);
    const double mu = atof(argv[8]), kappa = atof(argv[9]);
    const char *out = argv[10];
    const double Qarg = (argc > 11) ? atof(argv[11]) : 0.0;
    const int iters = (argc > 12) ? atoi(argv[12]) : 20000;
    double dtau = (argc > 13) ? atof(argv[13]) : 0.0;
    const char *thinit = (argc > 14) ? argv[14] : "bvp";
    const int printev = (argc > 15) ? atoi(argv[15]) : 500;
    const int wind = (argc > 16) ? atoi(argv[16]) : 0;
    const double wtor = (argc > 17) ? atof(argv[17]) : 2.0;
    const int twist = (argc > 18) ? atoi(argv[18]) : 0;
    const int fixJ = (argc > 19) ? atoi(argv[19]) : 0;
    const double Jarg = (argc > 20) ? atof(argv[20]) : 0.0;
    if (N < 4 || printev < 1 || iters < 0) {
        fprintf(stderr, "ERROR: N >= 4, iters >= 0 and printev >= 1 required\n");
        return 1;
    }
    const long N3 = (long)N * N * N;
    const double dx = 2.0 * L / (N - 1), dx2 = dx * dx;
    if (dtau <= 0) dtau = 0.04 * dx2;

    scp_qflow_seed_args sa = {N, L, omega0, eta, mth, !strcmp(thinit, "bvp"), wind, twist, wtor};
    double *st = malloc(sizeof(double) * 12 * (size_t)N3);
    if (!st) { fprintf(stderr, "FATAL: malloc failed\n"); return 1; }
    double gmax = 0;
    char err[600];
    if (scp_qflow_seed(&sa, prof, st, &gmax, err, sizeof err)) { fprintf(stderr, "%s\n", err); return 1; }

    scp_qflow_params qp;
    memset(&qp, 0, sizeof qp);
    qp.N = N; qp.L = L; qp.eta = eta; qp.m = m; qp.mth = mth; qp.mu = mu; qp.kappa = kappa;
    scp_qflow *f = NULL;
    if (scp_qflow_create(&qp, &f)) die_abi("qflow create");
    const double *cols[12];
    for (int q = 0; q < 12; q++) cols[q] = st + (size_t)q * N3;
    if (scp_qflow_upload(f, cols, 12)) die_abi("qflow upload");
    double nphi, nth;
    if (scp_qflow_norm(f, &nphi, &nth)) die_abi("qflow norm");
    const double Nnorm = nphi + nth;
    const double Q = (Qarg > 0) ? Qarg : omega0 * Nnorm;
    const double Jt = (fixJ && Jarg == 0.0 && wind > 0) ? (double)wind * Q : Jarg;
    printf("# eta_qflow N=%d L=%g dx=%.4f eta=%.3f m=%.3f mth=%.3f mu=%g kappa=%g\n", N, L, dx, eta, m, mth, mu,
           kappa);
    printf("# thinit=%s g_max=%.4e  N_init=%.4f  Q=%.4f  omega_init=Q/N=%.6f  dtau=%.5g iters=%d\n", thinit, gmax,
           Nnorm, Q, Q / Nnorm, dtau, iters);
    if (fixJ) {
        printf("# fixJ: J_target=%.4f  (two multipliers w,Om; Jhat = d_phi - zhat x)\n", Jt);
        printf("# %-8s %-10s %-10s %-12s %-12s %-12s %-12s %-10s %-10s\n", "iter", "omega", "Omega", "E_QJ", "maxF",
               "s_max", "N", "J_act", "cond");
        if (argc <= 13 || atof(argv[13]) <= 0) dtau = 0.02 * dx2;   /* after the header, as the reference */
    } else
        printf("# %-8s %-10s %-12s %-12s %-12s %-12s\n", "iter", "omega", "E_Q", "maxF", "s_max", "N");
    fflush(stdout);

    const int maxrows = iters / printev + 2;
    int nrows = 0;
    double omega = 0, last[5] = {0, 0, 1.0, 0, 0};
    if (fixJ) {
        scp_qflow_row_j *rows = calloc((size_t)maxrows, sizeof *rows);
        if (scp_qflow_run_fixj(f, Q, Jt, dtau, iters, printev, rows, maxrows, &nrows, last)) die_abi("qflow run");
        for (int r = 0; r < nrows && r < maxrows; r++)
            printf("  %-8d %-10.6f %-10.6f %-12.6e %-12.4e %-12.4e %-12.4f %-10.4f %-10.3e\n", rows[r].iter,
                   rows[r].omega, rows[r].Omega, rows[r].E_QJ, rows[r].maxF, rows[r].s_max, rows[r].N,
                   rows[r].J_act, rows[r].cond);
        free(rows);
        omega = last[0];
    } else {
        scp_qflow_row *rows = calloc((size_t)maxrows, sizeof *rows);
        if (scp_qflow_run(f, Q, dtau, iters, printev, rows, maxrows, &nrows, &omega)) die_abi("qflow run");
        for (int r = 0; r < nrows && r < maxrows; r++)
            printf("  %-8d %-10.6f %-12.6e %-12.4e %-12.4e %-12.4f\n", rows[r].iter, rows[r].omega, rows[r].E_Q,
                   rows[r].maxF, rows[r].s_max, rows[r].N);
        free(rows);
    }
    const double Omega = last[1], cond = last[2], ell = last[3], Kk = last[4];

    double Nphi, Nth;
    if (scp_qflow_norm(f, &Nphi, &Nth)) die_abi("qflow norm");
    double *fin[12];
    for (int q = 0; q < 12; q++) fin[q] = st + (size_t)q * N3;
    if (scp_qflow_download(f, fin, 12)) die_abi("qflow download");
    double *jh = NULL;                       /* Jhat of the final fields (fixJ) */
    if (fixJ) {
        jh = malloc(sizeof(double) * 12 * (size_t)N3);
        if (!jh) { fprintf(stderr, "FATAL: malloc failed\n"); return 1; }
        double *jc[12];
        for (int q = 0; q < 12; q++) jc[q] = jh + (size_t)q * N3;
        if (scp_qflow_download_jhat(f, jc, 12)) die_abi("qflow Jhat download");
    }
    scp_qflow_destroy(f);
    printf("# final: omega=%.6f  Q_phi=%.4f  Q_theta=%.4f  Q_total=%.4f (target %.4f)\n", omega, omega * Nphi,
           omega * Nth, omega * (Nphi + Nth), Q);
    if (fixJ) {
        const double Nfin = Nphi + Nth;
        printf("# fixJ final: Omega=%.6f  J_actual=%.4f (target %.4f)  Q_actual=%.4f  cond=%.3e  w_eff(Q/N)=%.6f\n",
               Omega, omega * ell - Omega * Kk, Jt, omega * Nfin - Omega * ell, cond, Q / Nfin);
    }
    if (omega >= m) printf("# WARNING: omega >= m (outside existence window)\n");
    if (omega * omega >= mth * mth) printf("# WARNING: omega >= m_theta (Theta unbound)\n");
    const double *u[3], *v[3], *tu[3], *tv[3];
    for (int a = 0; a < 3; a++) {
        u[a] = fin[a]; v[a] = fin[3 + a]; tu[a] = fin[6 + a]; tv[a] = fin[9 + a];
    }
    if (wind > 0 || twist != 0) {   /* the ring report of the final state (eta_qflow.c:404-432) */
        const long NN = (long)N * N;
        double Limz = 0, rho_pk = 0, amp_pk = 0;
        const int kz = N / 2;
        for (int i = 1; i < N - 1; i++) {
            const double x = -L + i * dx;
            for (int j = 1; j < N - 1; j++) {
                const double y = -L + j * dx;
                const long ii = (long)i * NN + (long)j * N + kz;
                double a2 = 0;
                for (int a = 0; a < 3; a++) a2 += u[a][ii] * u[a][ii] + v[a][ii] * v[a][ii];
                const double rho = sqrt(x * x + y * y);
                if (a2 > amp_pk) { amp_pk = a2; rho_pk = rho; }
            }
        }
        for (long ii = 0; ii < N3; ii++) {
            const int i = (int)(ii / NN), j = (int)((ii / N) % N), k = (int)(ii % N);
            if (i < 1 || i >= N - 1 || j < 1 || j >= N - 1 || k < 1 || k >= N - 1) continue;
            const double x = -L + i * dx, y = -L + j * dx;
            const long Xp = ii + NN, Xm = ii - NN, Yp = ii + N, Ym = ii - N;
            for (int a = 0; a < 3; a++) {
                const double dyu = (u[a][Yp] - u[a][Ym]) / (2 * dx), dyv = (v[a][Yp] - v[a][Ym]) / (2 * dx);
                const double dxu = (u[a][Xp] - u[a][Xm]) / (2 * dx), dxv = (v[a][Xp] - v[a][Xm]) / (2 * dx);
                const double dphiu = x * dyu - y * dxu, dphiv = x * dyv - y * dxv;
                Limz += u[a][ii] * dphiv - v[a][ii] * dphiu;
            }
        }
        const double Lz = omega * Limz * (dx * dx * dx);
        const int M = 256;
        double wsum = 0, prev = 0;
        for (int q = 0; q <= M; q++) {
            const double ph = 2.0 * M_PI * q / M;
            const double x = rho_pk * cos(ph), y = rho_pk * sin(ph);
            const int i = (int)lround((x + L) / dx), j = (int)lround((y + L) / dx);
            if (i < 0 || i >= N || j < 0 || j >= N) continue;
            const long ii = (long)i * NN + (long)j * N + kz;
            const double arg = atan2(v[0][ii], u[0][ii]);
            if (q > 0) {
                double d = arg - prev;
                while (d > M_PI) d -= 2 * M_PI;
                while (d < -M_PI) d += 2 * M_PI;
                wsum += d;
            }
            prev = arg;
        }
        printf("# ring: L_z=%.4f  L_z/Q=%.4f (target wind=%d)  rho_peak=%.3f  winding@peak=%.2f\n", Lz, Lz / Q, wind,
               rho_pk, wsum / (2 * M_PI));
    }

    /* the rotating seed, gen_qball_boost column order: Phidot = i omega phi, and
     * with fixJ - Omega Jhat phi (masked at the light cylinder, eta_qflow.c:213-217) */
    float *fc = malloc(sizeof(float) * 24 * (size_t)N3);
    if (!fc) { fprintf(stderr, "FATAL: malloc failed\n"); return 1; }
    for (long ii = 0; ii < N3; ii++) {
        double wm = 0;
        if (fixJ) {
            const int i = (int)(ii / ((long)N * N)), j = (int)((ii / N) % N);
            const double x = -L + i * dx, y = -L + j * dx, rho = sqrt(x * x + y * y);
            wm = 1.0 / (1.0 + exp((rho - 11.0) / 0.5));
        }
        for (int a = 0; a < 3; a++) {
            const double Ju = fixJ ? wm * jh[(size_t)a * N3 + ii] : 0, Jvv = fixJ ? wm * jh[(size_t)(3 + a) * N3 + ii] : 0;
            const double Jtu = fixJ ? wm * jh[(size_t)(6 + a) * N3 + ii] : 0;
            const double Jtv = fixJ ? wm * jh[(size_t)(9 + a) * N3 + ii] : 0;
            fc[(size_t)(0 + a) * N3 + ii] = (float)u[a][ii];
            fc[(size_t)(12 + a) * N3 + ii] = (float)v[a][ii];
            fc[(size_t)(6 + a) * N3 + ii] = (float)(-omega * v[a][ii] - Omega * Ju);
            fc[(size_t)(18 + a) * N3 + ii] = (float)(omega * u[a][ii] - Omega * Jvv);
            fc[(size_t)(3 + a) * N3 + ii] = (float)tu[a][ii];
            fc[(size_t)(15 + a) * N3 + ii] = (float)tv[a][ii];
            fc[(size_t)(9 + a) * N3 + ii] = (float)(-omega * tv[a][ii] - Omega * Jtu);
            fc[(size_t)(21 + a) * N3 + ii] = (float)(omega * tu[a][ii] - Omega * Jtv);
        }
    }
    free(jh);
    static const char *nm[24] = {"phi_x", "phi_y", "phi_z", "theta_x", "theta_y", "theta_z",
                                 "phi_vx", "phi_vy", "phi_vz", "theta_vx", "theta_vy", "theta_vz",
                                 "phiim_x", "phiim_y", "phiim_z", "thetaim_x", "thetaim_y", "thetaim_z",
                                 "phiim_vx", "phiim_vy", "phiim_vz", "thetaim_vx", "thetaim_vy", "thetaim_vz"};
    static const int sem[24] = {0, 0, 0, 1, 1, 1, 2, 2, 2, 2, 2, 2, 0, 0, 0, 1, 1, 1, 2, 2, 2, 2, 2, 2};
    static const int cmp[24] = {0, 1, 2, 0, 1, 2, 0, 1, 2, 3, 4, 5, 3, 4, 5, 3, 4, 5, 6, 7, 8, 9, 10, 11};
    sfaio *s = sfaio_create(out, N, N, N, L, L, L, 0.025 * dx);
    if (!s) { fprintf(stderr, "ERROR: cannot create output file '%s'\n", out); return 1; }
    /* the reference container's default codec (BSS + zstd, streaming); RAW without libzstd */
    if (sfaio_set_codec(s, sfaio_zstd_available() ? SFAIO_CODEC_BSS : SFAIO_CODEC_RAW, 1)) {
        fprintf(stderr, "FATAL: cannot select the codec\n");
        return 1;
    }
    for (int c = 0; c < 24; c++) sfaio_add_column(s, nm[c], SFAIO_F32, sem[c], cmp[c]);
    sfaio_finalize_header(s);
    const void *pc[24];
    for (int c = 0; c < 24; c++) pc[c] = fc + (size_t)c * N3;
    if (sfaio_write_frame(s, 0.0, pc)) { fprintf(stderr, "FATAL: frame write failed\n"); return 1; }
    sfaio_close(s);
    printf("eta_qflow: wrote %s\n", out);
    free(fc);
    free(st);
    return 0;
}
