/* seeds.c — deterministic Gaussian-ball initial states (see seeds.h). */
#include "seeds.h"

#include <math.h>
#include <string.h>

#define TWO_PI 6.283185307179586476925286766559
#define CUTOFF_SIGMAS 10.0    /* exp(-50) ~ 2e-22: contributions beyond are dropped */

static inline uint64_t mix64(uint64_t z) {          /* splitmix64 finaliser */
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

double scp_rng_u01(uint64_t seed, uint64_t stream, uint64_t index) {
    uint64_t h = mix64(mix64(seed) ^ (stream * 0xD1B54A32D192ED03ull));
    h = mix64(h ^ index);
    return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

double scp_rng_normal(uint64_t seed, uint64_t stream, uint64_t index) {   /* Box-Muller */
    double u1 = 1.0 - scp_rng_u01(seed, 2 * stream, index);               /* (0,1] */
    double u2 = scp_rng_u01(seed, 2 * stream + 1, index);
    return sqrt(-2.0 * log(u1)) * cos(TWO_PI * u2);
}

static inline double min_image(double d, double box) {
    d = fmod(d, box);
    if (d > 0.5 * box) d -= box;
    else if (d < -0.5 * box) d += box;
    return d;
}

int scp_seed_balls(int N, double L, int mode, const scp_ball *balls, int nballs,
                   double theta_sigma, double link_sigma, uint64_t seed, double *state) {
    if (N < 2 || (mode != 1 && mode != 2) || nballs < 0 || !state) return -1;
    const long n3 = (long)N * N * N, NN = (long)N * N;
    const int ncol = mode == 2 ? 30 : 24;
    const double dx = 2.0 * L / (N - 1), box = N * dx;
    #pragma omp parallel for schedule(static)
    for (long c = 0; c < (long)ncol * n3; c++) state[c] = 0.0;
    double *u[3], *tu[3], *ud[3], *v[3], *tv[3], *vd[3];
    for (int a = 0; a < 3; a++) {
        u[a] = state + (0 + a) * n3;   tu[a] = state + (3 + a) * n3;  ud[a] = state + (6 + a) * n3;
        v[a] = state + (12 + a) * n3;  tv[a] = state + (15 + a) * n3; vd[a] = state + (18 + a) * n3;
    }
    /* balls one at a time over their bounding boxes (periodic), so overlapping
     * balls never race; the site loop inside a ball is parallel */
    for (int b = 0; b < nballs; b++) {
        const scp_ball *B = &balls[b];
        int half = (int)ceil(CUTOFF_SIGMAS * B->sigma / dx);
        const int whole = 2 * half + 1 >= N;          /* box covers the grid: visit every site once */
        const int span = whole ? N : 2 * half + 1;
        if (whole) half = 0;
        const int ci = whole ? 0 : (int)lround((B->x0 + L) / dx);
        const int cj = whole ? 0 : (int)lround((B->y0 + L) / dx);
        const int ck = whole ? 0 : (int)lround((B->z0 + L) / dx);
        double cre[3], cim[3];
        for (int a = 0; a < 3; a++) {
            double ph = B->phase + B->winding * TWO_PI * a / 3.0;
            cre[a] = cos(ph); cim[a] = sin(ph);
        }
        #pragma omp parallel for schedule(static)
        for (int di = 0; di < span; di++) {
            int i = ((ci - half + di) % N + N) % N;
            double x = -L + i * dx, ex = min_image(x - B->x0, box);
            for (int dj = 0; dj < span; dj++) {
                int j = ((cj - half + dj) % N + N) % N;
                double y = -L + j * dx, ey = min_image(y - B->y0, box);
                for (int dk = 0; dk < span; dk++) {
                    int k = ((ck - half + dk) % N + N) % N;
                    double z = -L + k * dx, ez = min_image(z - B->z0, box);
                    long c = (long)i * NN + (long)j * N + k;
                    double r2 = ex * ex + ey * ey + ez * ez;
                    double env = B->amp * exp(-r2 / (2.0 * B->sigma * B->sigma));
                    for (int a = 0; a < 3; a++) {
                        double re = env * cre[a], im = env * cim[a];
                        u[a][c] += re; v[a][c] += im;
                        ud[a][c] += -B->omega * im;     /* d/dt (re + i im) = i omega (re + i im) */
                        vd[a][c] += B->omega * re;
                    }
                }
            }
        }
    }
    #pragma omp parallel for schedule(static)
    for (long c = 0; c < n3; c++) {
        for (int a = 0; a < 3; a++) {
            tu[a][c] = theta_sigma * scp_rng_normal(seed, (uint64_t)a, (uint64_t)c);
            tv[a][c] = theta_sigma * scp_rng_normal(seed, (uint64_t)(3 + a), (uint64_t)c);
        }
        if (mode == 2)
            for (int d = 0; d < 3; d++)
                state[(24 + d) * n3 + c] = link_sigma == 0.0 ? 0.0
                    : link_sigma * scp_rng_normal(seed, (uint64_t)(6 + d), (uint64_t)c);
    }
    return 0;
}

int scp_draw_balls(int N, double L, int nballs, uint64_t seed,
                   double a_lo, double a_hi, double w_lo, double w_hi,
                   double o_lo, double o_hi, int alternate_signs, scp_ball *out) {
    const double dx = 2.0 * L / (N - 1), box = N * dx;
    for (int b = 0; b < nballs; b++) {
        uint64_t s = 1000 + (uint64_t)b;
        scp_ball *B = &out[b];
        B->x0 = -L + box * scp_rng_u01(seed, s, 0);
        B->y0 = -L + box * scp_rng_u01(seed, s, 1);
        B->z0 = -L + box * scp_rng_u01(seed, s, 2);
        B->amp = a_lo + (a_hi - a_lo) * scp_rng_u01(seed, s, 3);
        B->sigma = (w_lo + (w_hi - w_lo) * scp_rng_u01(seed, s, 4)) * dx;
        double om = o_lo + (o_hi - o_lo) * scp_rng_u01(seed, s, 5);
        int sign = (alternate_signs && (b & 1)) ? -1 : 1;
        B->omega = sign * om;
        B->winding = sign;
        B->phase = TWO_PI * scp_rng_u01(seed, s, 6);
    }
    return nballs;
}
