/* qflow_seed.h — initial state of the fixed-charge gradient flow (the seed part
 * of sfa/seed/eta_qflow.c:104-190), on the host, FP64.
 *
 * The state is 12 arrays of N^3 doubles, site i*N*N + j*N + k:
 *   0-2 u_a (Re Phi_a), 3-5 v_a (Im Phi_a), 6-8 tu_a (Re Theta_a), 9-11 tv_a.
 */
#ifndef QFLOW_SEED_H
#define QFLOW_SEED_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int    N;
    double L;
    double omega0, eta, mth;    /* the Theta partner's BVP uses mth^2 - omega0^2 */
    int    theta_bvp;           /* 1: Theta from the BVP partner ("bvp"), 0: zero */
    int    wind, twist;         /* azimuthal / poloidal winding (0, 0: the ball) */
    double wtor;                /* major radius of the ring seed */
} scp_qflow_seed_args;

/* Reads the radial profile (r f), uniformly spaced from r = 0, and fills
 * state (12 N^3 doubles). *gmax: max |g| of the Theta partner. Returns 0, or
 * -1 with the reference's message in err. */
int scp_qflow_seed(const scp_qflow_seed_args *a, const char *profile, double *state, double *gmax, char *err,
                   int errlen);

#ifdef __cplusplus
}
#endif
#endif
