/* scp_init.c — initial states of the drop-in binary (see scp_init.h). */
#define _GNU_SOURCE
#include "scp_init.h"

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "sfa_io.h"

#define PI_ 3.14159265358979323846

static int ferr(char *err, int n, const char *fmt, const char *a) {
    snprintf(err, (size_t)n, fmt, a);
    return -1;
}

static int ncols_of(int mode) { return mode == 2 ? 30 : (mode == 1 ? 24 : 12); }

static void grid_geom(const scp_config *c, long *n3, double *dx) {
    *n3 = (long)c->N * c->N * c->N;
    *dx = 2.0 * c->L / (c->N - 1);
}

/* scp_init.h:19-32 */
static void init_oscillon(const scp_config *c, double *st) {
    long n3; double dx; grid_geom(c, &n3, &dx);
    const int N = c->N;
    printf("Init: oscillon (A=%.3f sigma=%.3f)\n", c->A, c->sigma);
    #pragma omp parallel for schedule(static)
    for (long q = 0; q < n3; q++) {
        int i = (int)(q / ((long)N * N)), j = (int)((q / N) % N), k = (int)(q % N);
        double x = -c->L + i * dx, y = -c->L + j * dx, z = -c->L + k * dx;
        double r2 = x * x + y * y + z * z;
        double env = c->A * exp(-r2 / (2.0 * c->sigma * c->sigma));
        for (int a = 0; a < 3; a++) st[a * n3 + q] = env * cos(c->delta[a]);
    }
}

/* scp_init.h:34-56 */
static void init_braid(const scp_config *c, double *st) {
    long n3; double dx; grid_geom(c, &n3, &dx);
    const int N = c->N;
    const double L = c->L, kw = PI_ / L, omega = sqrt(kw * kw + c->m2);
    const double sx = 1 + c->ellip, sy = 1 - c->ellip, inv2R2 = 1.0 / (2 * c->R_tube * c->R_tube);
    const double kb = PI_ / L, ob = sqrt(kb * kb + c->m2);
    printf("Init: braid (R=%.1f ellip=%.4f A=%.2f A_bg=%.2f)\n", c->R_tube, c->ellip, c->A, c->A_bg);
    #pragma omp parallel for schedule(static)
    for (long q = 0; q < n3; q++) {
        int i = (int)(q / ((long)N * N)), j = (int)((q / N) % N), k = (int)(q % N);
        double x = -L + i * dx, y = -L + j * dx, z = -L + k * dx;
        double env = exp(-(x * x / (sx * sx) + y * y / (sy * sy)) * inv2R2);
        for (int a = 0; a < 3; a++) {
            double ph = kw * z + c->delta[a], pb = kb * z + 2 * PI_ * a / 3.0;
            st[a * n3 + q] = c->A * env * cos(ph) + c->A_bg * cos(pb);
            st[(6 + a) * n3 + q] = omega * c->A * env * sin(ph) + ob * c->A_bg * sin(pb);
        }
    }
}

/* state column of an archive column (scp_init.h:92-126); -1 = not loaded */
static int target_col(int sem, int comp, int mode, int *warned_im, int *warned_gauge) {
    if (sem == SFAIO_POSITION && comp < 3) return comp;
    if (sem == SFAIO_ANGLE && comp < 3) return 3 + comp;
    if (sem == SFAIO_VELOCITY && comp < 3) return 6 + comp;
    if (sem == SFAIO_VELOCITY && comp >= 3 && comp < 6) return 6 + comp;
    int im = (sem == SFAIO_POSITION && comp >= 3 && comp < 6) || (sem == SFAIO_ANGLE && comp >= 3 && comp < 6)
          || (sem == SFAIO_VELOCITY && comp >= 6 && comp < 12);
    if (im) {
        if (mode >= 1) return sem == SFAIO_POSITION ? 9 + comp : 12 + comp;
        if (!*warned_im) { printf("  WARNING: file has imaginary-sector columns; skipped (complex_phi=0)\n"); *warned_im = 1; }
        return -1;
    }
    int ga = (sem == SFAIO_ANGLE && comp >= 6 && comp < 9) || (sem == SFAIO_VELOCITY && comp >= 12 && comp < 15);
    if (ga) {
        if (mode == 2) return sem == SFAIO_ANGLE ? 18 + comp : 15 + comp;
        if (!*warned_gauge) { printf("  WARNING: file has gauge-sector columns; skipped (complex_gauge=0)\n"); *warned_gauge = 1; }
    }
    return -1;
}

static void convert(const uint8_t *src, int dtype, double *dst, long n) {
    if (dtype == SFAIO_F64) memcpy(dst, src, (size_t)n * 8);
    else if (dtype == SFAIO_F32) { const float *f = (const float *)src; for (long i = 0; i < n; i++) dst[i] = (double)f[i]; }
    else if (dtype == SFAIO_F16) { const uint16_t *h = (const uint16_t *)src; for (long i = 0; i < n; i++) dst[i] = scp_f16_to_f64(h[i]); }
}

/* scp_init.h:58-161 */
/* Index of the k-th voxel frame from the end (k = 1: the last), -1 if there is
 * none. Vector frames (FRVD) hold fitted patches of six fields and no
 * velocities, so a restart or template reads the voxel frames only. */
static int voxel_frame_from_end(sfaio *s, int k) {
    for (int f = (int)s->total_frames - 1; f >= 0; f--)
        if (sfaio_frame_type(s, (uint32_t)f) == SFAIO_FRAME_VOXEL && --k == 0) return f;
    return -1;
}

static int init_from_sfa(const scp_config *c, int mode, const char *path, int frame, double *st, char *err, int n) {
    printf("Init: SFA file '%s' frame=%d\n", path, frame);
    sfaio *s = sfaio_open(path);
    if (!s) return ferr(err, n, "FATAL: cannot open SFA '%s'", path);
    long n3 = (long)c->N * c->N * c->N;
    if ((int)s->Nx != c->N || (int)s->Ny != c->N || (int)s->Nz != c->N) {
        snprintf(err, (size_t)n, "FATAL: SFA grid %ux%ux%u != sim grid %d^3", s->Nx, s->Ny, s->Nz, c->N);
        sfaio_close(s);
        return -1;
    }
    if (frame < 0) frame = voxel_frame_from_end(s, -frame);
    if (frame < 0 || frame >= (int)s->total_frames) {
        snprintf(err, (size_t)n, "FATAL: frame %d out of range [0,%u)", frame, s->total_frames);
        sfaio_close(s);
        return -1;
    }
    printf("  Grid: %ux%ux%u, %d columns, %u frames, reading frame %d\n", s->Nx, s->Ny, s->Nz,
           s->n_columns, s->total_frames, frame);
    uint8_t *buf = malloc(s->frame_bytes);
    int rc = buf ? sfaio_read_frame(s, (uint32_t)frame, buf) : -1;
    if (rc) {
        free(buf);
        sfaio_close(s);
        return ferr(err, n, "FATAL: cannot read frame of '%s'", path);
    }
    int loaded[30] = {0}, wim = 0, wg = 0;
    uint64_t off = 0;
    for (uint32_t col = 0; col < s->n_columns; col++) {
        const sfaio_col *k = &s->cols[col];
        int t = target_col(k->semantic, k->component, mode, &wim, &wg);
        if (t >= 0) {
            convert(buf + off, k->dtype, st + (long)t * n3, n3);
            loaded[t] = 1;
            printf("  Loaded col %u '%.12s' (sem=%d comp=%d dtype=%d) -> slot %d\n", col, k->name,
                   k->semantic, k->component, k->dtype, t);
        }
        off += (uint64_t)n3 * sfaio_dtype_size[k->dtype];
    }
    int nf = 0, nv = 0, ng = 0;
    for (int q = 0; q < 6; q++) nf += loaded[q];
    for (int q = 6; q < 12; q++) nv += loaded[q];
    for (int q = 24; q < 30; q++) ng += loaded[q];
    printf("  Loaded: %d/6 field arrays, %d/6 velocity arrays\n", nf, nv);
    if (nv == 0) printf("  WARNING: no velocity data — starting from rest (cold restart)\n");
    if (mode == 2 && c->g_gauge != 0 && ng == 0)
        printf("  WARNING: init=sfa: no gauge columns; E seeded by Gauss projection\n");
    if (fabs(s->Lx - c->L) > 0.01) printf("  WARNING: SFA L=%.2f != config L=%.2f\n", s->Lx, c->L);
    free(buf);
    sfaio_close(s);
    return 0;
}

/* scp_init.h:163-177 */
static int init_from_exec(const scp_config *c, int mode, double *st, char *err, int n) {
    printf("Init: exec '%s'\n", c->init_exec);
    char tmp[512], cmd[2048];
    snprintf(tmp, sizeof tmp, "/tmp/scp_sim_init_%d.sfa", (int)getpid());
    snprintf(cmd, sizeof cmd, "%s > %s", c->init_exec, tmp);
    printf("  Running: %s\n", cmd);
    int ret = system(cmd);
    if (ret != 0) { snprintf(err, (size_t)n, "FATAL: exec init failed (exit %d)", ret); return -1; }
    int rc = init_from_sfa(c, mode, tmp, -1, st, err, n);
    remove(tmp);
    return rc;
}

/* scp_init.h:179-271: analytic background plus a stamped template */
static int init_template(const scp_config *c, double *st, char *err, int n) {
    printf("Init: template '%s'\n", c->init_sfa);
    long n3; double dx; grid_geom(c, &n3, &dx);
    const int N = c->N;
    const double L = c->L, kb = PI_ / L, ob = sqrt(kb * kb + c->m2);
    for (long q = 0; q < n3; q++) {
        int i = (int)(q / ((long)N * N)), k = (int)(q % N);
        double x = -L + i * dx, z = -L + k * dx, A = c->A_bg;
        if (c->bc_type == 1) {
            double fr = (x + L) / (2.0 * L);
            A = c->gradient_A_high * (1.0 - fr) + c->gradient_A_low * fr;
        }
        for (int a = 0; a < 3; a++) {
            double pb = kb * z + 2 * PI_ * a / 3.0;
            st[a * n3 + q] = A * cos(pb);
            st[(6 + a) * n3 + q] = ob * A * sin(pb);
        }
    }
    sfaio *s = sfaio_open(c->init_sfa);
    if (!s) return ferr(err, n, "FATAL: cannot open template '%s'", c->init_sfa);
    const int TN = (int)s->Nx;
    const double TL = s->Lx, Tdx = 2.0 * TL / (TN - 1);
    const long T3 = (long)TN * TN * TN;
    printf("  Template: %d^3, L=%.2f, dx=%.4f\n", TN, TL, Tdx);
    uint8_t *buf = malloc(s->frame_bytes);
    double *tv = calloc((size_t)12 * T3, sizeof(double));
    const int tf = voxel_frame_from_end(s, 1);
    if (!buf || !tv || tf < 0 || sfaio_read_frame(s, (uint32_t)tf, buf)) {
        free(buf); free(tv); sfaio_close(s);
        return ferr(err, n, "FATAL: cannot read template '%s'", c->init_sfa);
    }
    uint64_t off = 0;
    for (uint32_t col = 0; col < s->n_columns; col++) {
        const sfaio_col *k = &s->cols[col];
        int t = -1;
        if (k->semantic == SFAIO_POSITION && k->component < 3) t = k->component;
        else if (k->semantic == SFAIO_ANGLE && k->component < 3) t = 3 + k->component;
        else if (k->semantic == SFAIO_VELOCITY && k->component < 6) t = 6 + k->component;
        if (t >= 0) convert(buf + off, k->dtype, tv + (long)t * T3, T3);
        off += (uint64_t)T3 * sfaio_dtype_size[k->dtype];
    }
    free(buf);
    sfaio_close(s);
    int placed = 0;
    for (int ti = 0; ti < TN; ti++) {
        double tx = -TL + ti * Tdx; int gi = (int)((tx + L) / dx + 0.5);
        if (gi < 0 || gi >= N) continue;
        for (int tj = 0; tj < TN; tj++) {
            double ty = -TL + tj * Tdx; int gj = (int)((ty + L) / dx + 0.5);
            if (gj < 0 || gj >= N) continue;
            for (int tk = 0; tk < TN; tk++) {
                double tz = -TL + tk * Tdx; int gk = (int)((tz + L) / dx + 0.5);
                if (gk < 0 || gk >= N) continue;
                long ts = ((long)ti * TN + tj) * TN + tk, gs = ((long)gi * N + gj) * N + gk;
                for (int a = 0; a < 3; a++) {
                    double pb = kb * tz + 2 * PI_ * a / 3.0;
                    st[a * n3 + gs] += tv[a * T3 + ts] - c->A_bg * cos(pb);
                    st[(6 + a) * n3 + gs] += tv[(6 + a) * T3 + ts] - ob * c->A_bg * sin(pb);
                    st[(3 + a) * n3 + gs] += tv[(3 + a) * T3 + ts];
                    st[(9 + a) * n3 + gs] += tv[(9 + a) * T3 + ts];
                }
                placed++;
            }
        }
    }
    free(tv);
    printf("  Placed %d voxels from template\n", placed);
    return 0;
}

/* linear interpolation on a strictly increasing table (scp_init.h:277-288) */
static double tab(const double *rt, const double *vt, size_t np, double r, double tail) {
    if (r <= rt[0]) return vt[0];
    if (r >= rt[np - 1]) return tail;
    size_t lo = 0, hi = np - 1;
    while (hi - lo > 1) { size_t m = (lo + hi) / 2; if (rt[m] <= r) lo = m; else hi = m; }
    double w = (r - rt[lo]) / (rt[hi] - rt[lo]);
    return vt[lo] + w * (vt[hi] - vt[lo]);
}

/* radial E with a linear core and an inverse-square tail (scp_init.h:292-306) */
static double tab_er(const double *rt, const double *et, size_t np, double r) {
    if (r <= rt[0]) return rt[0] > 0 ? et[0] * (r / rt[0]) : et[0];
    if (r >= rt[np - 1]) { double R = rt[np - 1]; return et[np - 1] * (R / r) * (R / r); }
    return tab(rt, et, np, r, 0.0);
}

/* scp_init.h:314-486 */
static int init_qball(const scp_config *c, int mode, double *st, char *err, int n) {
    FILE *pf = fopen(c->qball_profile, "r");
    if (!pf) return ferr(err, n, "FATAL: cannot open qball profile '%s'", c->qball_profile);
    size_t cap = 256, np = 0;
    int ncol = 0, lineno = 0;
    double *rt = malloc(cap * 8), *ft = malloc(cap * 8), *et = malloc(cap * 8), *wt = malloc(cap * 8);
    char line[1024];
    int bad = 0;
    while (!bad && fgets(line, sizeof line, pf)) {
        lineno++;
        char *p = line;
        while (*p == ' ' || *p == '\t') p++;
        if (*p == '#' || *p == '\n' || *p == '\r' || *p == '\0') continue;
        double r, f, er = 0, we = 0;
        int nf = sscanf(p, "%lf %lf %lf %lf", &r, &f, &er, &we);
        if (nf < 2) { snprintf(err, (size_t)n, "FATAL: qball profile parse error at %s:%d", c->qball_profile, lineno); bad = 1; break; }
        if (ncol == 0) ncol = nf;
        else if (nf < ncol) { snprintf(err, (size_t)n, "FATAL: qball profile column count changed at %s:%d (%d < %d)", c->qball_profile, lineno, nf, ncol); bad = 1; break; }
        if (np == 0 && r < 0) { snprintf(err, (size_t)n, "FATAL: qball profile first r must be >= 0 (got %f)", r); bad = 1; break; }
        if (np > 0 && r <= rt[np - 1]) { snprintf(err, (size_t)n, "FATAL: qball profile r not strictly increasing at %s:%d", c->qball_profile, lineno); bad = 1; break; }
        if (np == cap) {
            cap *= 2;
            rt = realloc(rt, cap * 8); ft = realloc(ft, cap * 8); et = realloc(et, cap * 8); wt = realloc(wt, cap * 8);
        }
        rt[np] = r; ft[np] = f; et[np] = er; wt[np] = we; np++;
    }
    fclose(pf);
    if (!bad && np < 2) { snprintf(err, (size_t)n, "FATAL: qball profile has <2 points"); bad = 1; }
    if (bad) { free(rt); free(ft); free(et); free(wt); return -1; }
    const double om = c->qball_omega, GG = c->g_gauge;
    const int gauged = (mode == 2 && GG != 0.0);
    int have_er = ncol >= 3;
    if (!gauged || ncol == 2) {
        for (size_t q = 0; q < np; q++) wt[q] = om;
        if (gauged && ncol == 2)
            printf("  WARNING: ungauged profile (no Er column): seeding vdot=omega*f, E from projection; "
                   "O(g^2) settling transient expected\n");
        have_er = gauged && have_er;
    } else if (ncol == 3) {                 /* weff = omega + g*a0(r), a0 from the Er column */
        double a0 = -(et[np - 1] * rt[np - 1]);
        wt[np - 1] = om + GG * a0;
        for (size_t q = np - 1; q > 0; q--) {
            a0 -= 0.5 * (et[q - 1] + et[q]) * (rt[q] - rt[q - 1]);
            wt[q - 1] = om + GG * a0;
        }
    }
    printf("Init: qball (%zu points, %d cols, r=[%.3f,%.3f], f0=%.6f, omega=%.4f, weff0=%.4f; delta[] ignored)\n",
           np, ncol, rt[0], rt[np - 1], ft[0], om, wt[0]);
    if (fabs(ft[np - 1]) > 1e-8)
        printf("  WARNING: profile truncated while still sizable (f_last=%.3e at r=%.3f)\n", ft[np - 1], rt[np - 1]);
    long n3; double dx; grid_geom(c, &n3, &dx);
    const int N = c->N;
    const double L = c->L;
    const double x0 = c->qball_x0 >= 1e29 ? 0.0 : c->qball_x0, y0 = c->qball_y0 >= 1e29 ? 0.0 : c->qball_y0;
    const double z0 = c->qball_z0 >= 1e29 ? 0.0 : c->qball_z0;
    const int ball2 = c->qball2_x0 < 1e29 && c->qball2_y0 < 1e29 && c->qball2_z0 < 1e29;
    const double x2 = ball2 ? c->qball2_x0 : 0, y2 = ball2 ? c->qball2_y0 : 0, z2 = ball2 ? c->qball2_z0 : 0;
    const double sg = (double)c->qball2_sign, cd = ball2 ? cos(c->qball2_phase) : 1.0, sd = ball2 ? sin(c->qball2_phase) : 0.0;
    if (ball2)
        printf("  ball1 center=(%.2f,%.2f,%.2f); ball2 center=(%.2f,%.2f,%.2f) sign=%+d phase=%.4f\n",
               x0, y0, z0, x2, y2, z2, c->qball2_sign, c->qball2_phase);
    #pragma omp parallel for schedule(static)
    for (long q = 0; q < n3; q++) {
        int i = (int)(q / ((long)N * N)), j = (int)((q / N) % N), k = (int)(q % N);
        double x = -L + i * dx, y = -L + j * dx, z = -L + k * dx;
        double r = sqrt((x - x0) * (x - x0) + (y - y0) * (y - y0) + (z - z0) * (z - z0));
        double f = tab(rt, ft, np, r, 0.0), we = tab(rt, wt, np, r, wt[np - 1]);
        for (int a = 0; a < 3; a++) {
            st[a * n3 + q] = f; st[(12 + a) * n3 + q] = 0.0;
            st[(6 + a) * n3 + q] = 0.0; st[(18 + a) * n3 + q] = we * f;
        }
        if (ball2) {
            double r2 = sqrt((x - x2) * (x - x2) + (y - y2) * (y - y2) + (z - z2) * (z - z2));
            double f2 = tab(rt, ft, np, r2, 0.0), we2 = tab(rt, wt, np, r2, wt[np - 1]);
            for (int a = 0; a < 3; a++) {
                st[a * n3 + q] += f2 * cd; st[(12 + a) * n3 + q] += f2 * sd;
                st[(6 + a) * n3 + q] += -sg * we2 * f2 * sd; st[(18 + a) * n3 + q] += sg * we2 * f2 * cd;
            }
        }
        if (gauged && have_er) {            /* E on radial link midpoints, scp_init.h:454-484 */
            for (int d = 0; d < 3; d++) {
                double mx = x + (d == 0 ? 0.5 * dx : 0) - x0, my = y + (d == 1 ? 0.5 * dx : 0) - y0;
                double mz = z + (d == 2 ? 0.5 * dx : 0) - z0, rm = sqrt(mx * mx + my * my + mz * mz);
                double E = 0.0;
                if (rm > 1e-12) E += tab_er(rt, et, np, rm) * ((d == 0 ? mx : d == 1 ? my : mz) / rm);
                if (ball2) {
                    double ax = x + (d == 0 ? 0.5 * dx : 0) - x2, ay = y + (d == 1 ? 0.5 * dx : 0) - y2;
                    double az = z + (d == 2 ? 0.5 * dx : 0) - z2, r2m = sqrt(ax * ax + ay * ay + az * az);
                    if (r2m > 1e-12) E += sg * tab_er(rt, et, np, r2m) * ((d == 0 ? ax : d == 1 ? ay : az) / r2m);
                }
                st[(27 + d) * n3 + q] += E;
            }
        }
    }
    free(rt); free(ft); free(et); free(wt);
    return 0;
}

int scp_do_init(const scp_config *c, int mode, double *st, char *err, int n) {
    long n3 = (long)c->N * c->N * c->N;
    memset(st, 0, sizeof(double) * (size_t)ncols_of(mode) * (size_t)n3);
    if (!strcmp(c->init, "oscillon")) { init_oscillon(c, st); return 0; }
    if (!strcmp(c->init, "braid")) { init_braid(c, st); return 0; }
    if (!strcmp(c->init, "sfa")) return init_from_sfa(c, mode, c->init_sfa, c->init_frame, st, err, n);
    if (!strcmp(c->init, "exec")) return init_from_exec(c, mode, st, err, n);
    if (!strcmp(c->init, "template")) return init_template(c, st, err, n);
    if (!strcmp(c->init, "qball") && mode >= 1) return init_qball(c, mode, st, err, n);
    return ferr(err, n, "FATAL: unknown init mode '%s'", c->init);
}

void scp_net_charge(const scp_config *c, int mode, const double *s, double *qn, double *qa) {
    long n3; double dx; grid_geom(c, &n3, &dx);
    const double dV = dx * dx * dx;
    double a = 0, b = 0;
    if (mode >= 1) {
        #pragma omp parallel for reduction(+:a,b) schedule(static)
        for (long q = 0; q < n3; q++) {
            double r = 0;
            for (int k = 0; k < 3; k++)
                r += s[k * n3 + q] * s[(18 + k) * n3 + q] - s[(12 + k) * n3 + q] * s[(6 + k) * n3 + q]
                   + s[(3 + k) * n3 + q] * s[(21 + k) * n3 + q] - s[(15 + k) * n3 + q] * s[(9 + k) * n3 + q];
            a += r * dV; b += fabs(r) * dV;
        }
    }
    *qn = a; *qa = b;
}

void scp_gauge_transform(const scp_config *c, double *s) {
    long n3; double dx; grid_geom(c, &n3, &dx);
    const int N = c->N;
    const double GG = c->g_gauge;
    double *al = malloc(sizeof(double) * (size_t)n3);
    for (long q = 0; q < n3; q++) {               /* counter-keyed LCG, scp_sim.c:2486-2488 */
        uint64_t g = (uint64_t)q * 6364136223846793005ULL + 1442695040888963407ULL;
        g = g * 6364136223846793005ULL + 1442695040888963407ULL;
        al[q] = (((double)(g >> 11)) / 9007199254740992.0 - 0.5) * 20.0;
    }
    static const int RE[4] = {0, 6, 3, 9}, IM[4] = {12, 18, 15, 21};
    #pragma omp parallel for schedule(static)
    for (long q = 0; q < n3; q++) {
        int i = (int)(q / ((long)N * N)), j = (int)((q / N) % N), k = (int)(q % N);
        long np[3] = {((long)((i + 1) % N) * N + j) * N + k, ((long)i * N + (j + 1) % N) * N + k,
                      ((long)i * N + j) * N + (k + 1) % N};
        double ca = cos(GG * al[q]), sa = sin(GG * al[q]);
        for (int p = 0; p < 4; p++)
            for (int a = 0; a < 3; a++) {
                double *re = s + (RE[p] + a) * n3 + q, *im = s + (IM[p] + a) * n3 + q;
                double u = *re, v = *im;
                *re = ca * u + sa * v; *im = ca * v - sa * u;
            }
        for (int d = 0; d < 3; d++) {
            double v = s[(24 + d) * n3 + q] + GG * (al[np[d]] - al[q]);
            s[(24 + d) * n3 + q] = v - 2.0 * PI_ * rint(v / (2.0 * PI_));
        }
    }
    free(al);
}
