/* scp_init.h — initial states of the drop-in binary: the reference init modes
 * (sfa/sim/scp_init.h:19-499) written into a host state buffer in snapshot
 * column order (12/24/30 columns, index i*N*N + j*N + k).
 *
 *   oscillon  scp_init.h:19-32     real part Phi_a = A exp(-r^2/2s^2) cos(delta_a)
 *   braid     scp_init.h:34-56     real mode only
 *   sfa       scp_init.h:58-161    columns mapped by (semantic, component); f16/f32/f64
 *   exec      scp_init.h:163-177   run a generator, load its SFA output
 *   template  scp_init.h:179-271   background + stamped template by physical coordinate
 *   qball     scp_init.h:314-486   radial profile table (r f [Er [weff]]), optional 2nd ball
 */
#ifndef SCP_INIT_H
#define SCP_INIT_H

#include "scp_config.h"

#ifdef __cplusplus
extern "C" {
#endif

/* mode: 0 real (12 cols), 1 complex (24), 2 gauged (30). Returns 0 or -1
 * with a message in err (the binary prints it and exits 1, like the reference). */
int scp_do_init(const scp_config *c, int mode, double *state, char *err, int errlen);

/* Noether-charge totals of a state (net_charge_check, scp_sim.c:2264-2279). */
void scp_net_charge(const scp_config *c, int mode, const double *state, double *q_net, double *q_abs);

/* The self-test's random lattice gauge transformation (scp_sim.c:2485-2514). */
void scp_gauge_transform(const scp_config *c, double *state);

#ifdef __cplusplus
}
#endif
#endif
