/* qflow_seed.c — the seed of the fixed-charge gradient flow (eta_qflow.c:104-190).
 *
 * Profile: lines "r f" (lines starting with '#' or empty are skipped), spacing
 * dr = r[1] - r[0] assumed uniform from r = 0 (eta_qflow.c:106-113). The Theta
 * partner g(r) solves g'' + (2/r) g' - (2/r^2) g - (mth^2 - w0^2) g = -sqrt(3) eta f'
 * on the profile's own radial grid by the tridiagonal (Thomas) elimination,
 * g = 0 at both ends (eta_qflow.c:115-135). The ball seeds u_a = f(r) (all three
 * components) and Theta = g(r) (rhat x uhat) with uhat = (1,1,1)/sqrt(3); the ring
 * seeds (wind > 0 or twist != 0) place the profile on the tube around the circle
 * rho = wtor with the winding phase and leave Theta zero (eta_qflow.c:151-183).
 * Linear interpolation on the uniform grid as the reference does it; every
 * expression keeps the reference's order so that the seed is the same to the bit.
 */
#include "qflow_seed.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static int serr(char *err, int n, const char *msg) {
    if (err && n > 0) snprintf(err, (size_t)n, "%s", msg);
    return -1;
}

/* value of table t (np entries, spacing dr from r = 0) at r < rmax */
static double lin(const double *t, int np, double dr, double r) {
    double q = r / dr;
    int id = (int)q;
    if (id >= np - 1) id = np - 2;
    double fr = q - id;
    return t[id] + fr * (t[id + 1] - t[id]);
}

int scp_qflow_seed(const scp_qflow_seed_args *a, const char *profile, double *st, double *gmax_out, char *err,
                   int errlen) {
    const int N = a->N;
    const long N3 = (long)N * N * N, NN = (long)N * N;
    const double L = a->L, dx = 2.0 * L / (N - 1);
    int cap = 8192, np = 0;
    double *pr = malloc(cap * sizeof(double)), *pf = malloc(cap * sizeof(double));
    FILE *fp = fopen(profile, "r");
    if (!fp) {
        free(pr); free(pf);
        char m[600];
        snprintf(m, sizeof m, "cannot open %s", profile);
        return serr(err, errlen, m);
    }
    char ln[512];
    while (fgets(ln, sizeof ln, fp)) {
        if (ln[0] == '#' || ln[0] == '\n') continue;
        double r, f;
        if (sscanf(ln, "%lf %lf", &r, &f) == 2) {
            if (np >= cap) {
                cap *= 2;
                pr = realloc(pr, cap * sizeof(double));
                pf = realloc(pf, cap * sizeof(double));
            }
            pr[np] = r; pf[np] = f; np++;
        }
    }
    fclose(fp);
    if (np < 4) { free(pr); free(pf); return serr(err, errlen, "profile too short"); }
    const double dr = pr[1] - pr[0], rmax = pr[np - 1];

    /* f' and the Theta partner g */
    double *df = malloc(np * sizeof(double)), *gg = calloc(np, sizeof(double));
    for (int i = 1; i < np - 1; i++) df[i] = (pf[i + 1] - pf[i - 1]) / (2 * dr);
    df[0] = (pf[1] - pf[0]) / dr;
    df[np - 1] = 0;
    if (a->theta_bvp && a->eta != 0.0) {
        const double k2 = a->mth * a->mth - a->omega0 * a->omega0;
        double *A = malloc(np * sizeof(double)), *B = malloc(np * sizeof(double));
        double *C = malloc(np * sizeof(double)), *D = malloc(np * sizeof(double));
        double *cp = malloc(np * sizeof(double)), *dp = malloc(np * sizeof(double));
        for (int i = 1; i < np - 1; i++) {
            const double r = pr[i];
            A[i] = 1.0 / (dr * dr) - 1.0 / (r * dr);
            B[i] = -2.0 / (dr * dr) - 2.0 / (r * r) - k2;
            C[i] = 1.0 / (dr * dr) + 1.0 / (r * dr);
            D[i] = -sqrt(3.0) * a->eta * df[i];
        }
        const int lo = 1, hi = np - 2;
        cp[lo] = C[lo] / B[lo];
        dp[lo] = D[lo] / B[lo];
        for (int i = lo + 1; i <= hi; i++) {
            const double mm = B[i] - A[i] * cp[i - 1];
            cp[i] = C[i] / mm;
            dp[i] = (D[i] - A[i] * dp[i - 1]) / mm;
        }
        gg[hi] = dp[hi];
        for (int i = hi - 1; i >= lo; i--) gg[i] = dp[i] - cp[i] * gg[i + 1];
        free(A); free(B); free(C); free(D); free(cp); free(dp);
    }
    double gmax = 0;
    for (int i = 0; i < np; i++)
        if (fabs(gg[i]) > gmax) gmax = fabs(gg[i]);

    memset(st, 0, sizeof(double) * 12 * (size_t)N3);
    double *u[3], *v[3], *tu[3];
    for (int c = 0; c < 3; c++) {
        u[c] = st + (size_t)c * N3;
        v[c] = st + (size_t)(3 + c) * N3;
        tu[c] = st + (size_t)(6 + c) * N3;
    }
    const double inv = 1.0 / sqrt(3.0), ux = inv, uy = inv, uz = inv;
    const int wind = a->wind, twist = a->twist;
    const double wtor = a->wtor;
#pragma omp parallel for
    for (long ii = 0; ii < N3; ii++) {
        const int i = (int)(ii / NN), j = (int)((ii / N) % N), k = (int)(ii % N);
        const double x = -L + i * dx, y = -L + j * dx, z = -L + k * dx, r = sqrt(x * x + y * y + z * z);
        double f = 0, g = 0;
        if (r < rmax) {
            f = lin(pf, np, dr, r);
            g = lin(gg, np, dr, r);
        }
        if (wind > 0 || twist != 0) {
            const double rho = sqrt(x * x + y * y), phi_az = atan2(y, x);
            const double drho = rho - wtor, d = sqrt(drho * drho + z * z);
            const double chi = atan2(z, drho);
            double fd = 0;
            if (d < rmax) fd = lin(pf, np, dr, d);
            double amp = fd;
            for (int q = 0; q < wind; q++) amp *= rho / sqrt(rho * rho + 1.0);
            for (int q = 0; q < abs(twist); q++) amp *= d / sqrt(d * d + 1.0);
            const double ph = wind * phi_az + twist * chi;
            const double cw = cos(ph), sw = sin(ph);
            for (int c = 0; c < 3; c++) { u[c][ii] = amp * cw; v[c][ii] = amp * sw; }
            continue;
        }
        for (int c = 0; c < 3; c++) u[c][ii] = f;
        const double rx = (r > 1e-9) ? x / r : 0, ry = (r > 1e-9) ? y / r : 0, rz = (r > 1e-9) ? z / r : 0;
        tu[0][ii] = g * (ry * uz - rz * uy);
        tu[1][ii] = g * (rz * ux - rx * uz);
        tu[2][ii] = g * (rx * uy - ry * ux);
    }
    free(pr); free(pf); free(df); free(gg);
    if (gmax_out) *gmax_out = gmax;
    return 0;
}
