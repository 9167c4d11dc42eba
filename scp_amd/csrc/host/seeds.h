/* seeds.h — deterministic initial states for the sfa/sim benchmark workloads.
 *
 * The reference has no built-in init that produces Gaussian Q-ball sets
 * (sfa/sim/scp_init.h:489-499); its users write seed SFA files with separate
 * generators (the sfa/seed directory) and load them with init=sfa. These routines are
 * this project's generator: they fill a state buffer in SFA column order
 * (12/24/30 columns, see include/scp_amd.h) that is then written to an SFA
 * file or uploaded directly.
 *
 * Randomness is counter-based (splitmix64 keyed by seed, stream, site), so a
 * value depends only on (seed, stream, site) and never on thread order.
 */
#ifndef SCP_SEEDS_H
#define SCP_SEEDS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    double x0, y0, z0;   /* centre, physical coordinates */
    double amp;          /* peak |Phi_a| */
    double sigma;        /* Gaussian width, physical units */
    double omega;        /* internal rotation: dPhi/dt = i*omega*Phi */
    double phase;        /* overall phase */
    int    winding;      /* +1: Phi_a ~ e^{+i 2 pi a/3}; -1: conjugate winding */
} scp_ball;

/* Uniform double in [0,1) and standard normal keyed by (seed, stream, index). */
double scp_rng_u01(uint64_t seed, uint64_t stream, uint64_t index);
double scp_rng_normal(uint64_t seed, uint64_t stream, uint64_t index);

/* Fill `state` (ncols*N^3 doubles, mode 1 complex -> 24 cols, mode 2 gauged
 * -> 30 cols) with a superposition of Gaussian balls on the periodic box of
 * side N*dx (minimum-image distances), Theta noise N(0, theta_sigma^2) on both
 * tu and tv, zero Theta velocities, link noise N(0, link_sigma^2) (gauged),
 * E = 0. Returns 0, or -1 for bad arguments. */
int scp_seed_balls(int N, double L, int mode, const scp_ball *balls, int nballs,
                   double theta_sigma, double link_sigma, uint64_t seed, double *state);

/* Draw the random ball set of the 512^3 / batched / serving workloads:
 * centres uniform in the box, amp U[a_lo,a_hi], width U[w_lo,w_hi] cells,
 * omega U[o_lo,o_hi], phase U[0,2pi). Signs alternate (+,-,+,...) so the
 * total Noether charge is balanced in pairs. Returns nballs. */
int scp_draw_balls(int N, double L, int nballs, uint64_t seed,
                   double a_lo, double a_hi, double w_lo, double w_hi,
                   double o_lo, double o_hi, int alternate_signs, scp_ball *out);

#ifdef __cplusplus
}
#endif
#endif
