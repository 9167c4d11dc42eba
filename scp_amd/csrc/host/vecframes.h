/* vecframes.h — vector frames (FRVD) of the real mode: per-patch tricubic
 * fits of the six fields, a per-coefficient temporal model, I-frames (all
 * coefficients and the model) and P-frames (the patches whose residual
 * against the model exceeds a tolerance), written into the run's SFA file
 * (scp_sim.c:2942-2988 set-up, 3098-3193 per frame; payload layouts
 * sfa.h:1643-1720).
 *
 * The fits come from the GPU (scp_grid_vec_fit); this module holds what the
 * reference does on the host around them, in FP32 with the same operations
 * in the same order, so frame selection and payloads follow it exactly. */
#pragma once

#include <stdint.h>

#include "sfa_io.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int bs, n_patches, nc;           /* block edge, patches, coefficients per patch (6 x 64) */
    int iframe_interval, refit_interval;
    double delta_tol;
    float omega;                     /* breathing frequency of the temporal model */
    int frame, count;                /* frames written; frames in the running sums */
    int16_t *origins;                /* n_patches x (i, j, k) */
    float *coeffs;                   /* n_patches x nc: the current fit (filled by the caller) */
    float *mean, *amp, *phase;       /* temporal model per coefficient */
    float *sum_mean, *sum_cos, *sum_sin;
    float *delta;                    /* P-frame residuals */
    uint32_t *idx;                   /* P-frame patch indices */
} scp_vecframes;

/* N: grid edge; bs: block edge (vec_block_size, 1..N); iframe_interval >= 1;
 * 0, or -1 on bad arguments or allocation failure. */
int  scp_vf_init(scp_vecframes *v, int N, int bs, int iframe_interval, double delta_tol);
/* Update the temporal model with v->coeffs at time t and write an I- or P-frame. */
int  scp_vf_write(scp_vecframes *v, sfaio *out, double t);
void scp_vf_free(scp_vecframes *v);

#ifdef __cplusplus
}
#endif
