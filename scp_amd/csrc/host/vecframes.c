/* vecframes.c — host side of the real-mode vector frames (see vecframes.h).
 * FP32 arithmetic with the reference's operations and order (scp_sim.c:
 * 3118-3191); this file is compiled without FMA contraction. */
#include "vecframes.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define VF_FIELDS 6
#define VF_NC1 64     /* tricubic: 4 x 4 x 4 coefficients per field */

int scp_vf_init(scp_vecframes *v, int N, int bs, int iframe_interval, double delta_tol) {
    memset(v, 0, sizeof *v);
    if (bs < 1 || bs > N || iframe_interval < 1) return -1;
    const int bn = N / bs;
    v->bs = bs;
    v->n_patches = bn * bn * bn;
    v->nc = VF_FIELDS * VF_NC1;
    v->iframe_interval = iframe_interval;
    v->refit_interval = 50;                          /* scp_sim.c:2962 */
    v->delta_tol = delta_tol;
    v->omega = 2.0f * 3.14159265f / 2.2f;            /* as the reference forms it, in FP32 */
    const long nt = (long)v->n_patches * v->nc;
    v->origins = malloc(sizeof(int16_t) * 3 * (size_t)(v->n_patches > 0 ? v->n_patches : 1));
    v->idx = malloc(sizeof(uint32_t) * (size_t)(v->n_patches > 0 ? v->n_patches : 1));
    float **arrs[] = {&v->coeffs, &v->mean, &v->amp, &v->phase, &v->sum_mean, &v->sum_cos, &v->sum_sin, &v->delta};
    for (size_t a = 0; a < sizeof arrs / sizeof arrs[0]; a++)
        *arrs[a] = calloc((size_t)(nt > 0 ? nt : 1), sizeof(float));
    if (!v->origins || !v->idx || !v->coeffs || !v->mean || !v->amp || !v->phase || !v->sum_mean ||
        !v->sum_cos || !v->sum_sin || !v->delta) {
        scp_vf_free(v);
        return -1;
    }
    int p = 0;                                       /* patch order: i, then j, then k */
    for (int bi = 0; bi < bn; bi++)
        for (int bj = 0; bj < bn; bj++)
            for (int bk = 0; bk < bn; bk++, p++) {
                v->origins[3 * p + 0] = (int16_t)(bi * bs);
                v->origins[3 * p + 1] = (int16_t)(bj * bs);
                v->origins[3 * p + 2] = (int16_t)(bk * bs);
            }
    return 0;
}

void scp_vf_free(scp_vecframes *v) {
    free(v->origins); free(v->idx); free(v->coeffs); free(v->mean); free(v->amp); free(v->phase);
    free(v->sum_mean); free(v->sum_cos); free(v->sum_sin); free(v->delta);
    memset(v, 0, sizeof *v);
}

static void put(uint8_t *b, uint64_t *o, const void *src, uint64_t n) {
    memcpy(b + *o, src, n);
    *o += n;
}

int scp_vf_write(scp_vecframes *v, sfaio *out, double t) {
    const long nt = (long)v->n_patches * v->nc;
    const int iframe = v->frame == 0 || v->frame % v->iframe_interval == 0;
    const float *c = v->coeffs;
    /* temporal model: running sums, bootstrap on the first frame, refit at I-frames */
    const float cw = cosf(v->omega * (float)t), sw = sinf(v->omega * (float)t);
    for (long i = 0; i < nt; i++) {
        v->sum_mean[i] += c[i];
        v->sum_cos[i] += c[i] * cw;
        v->sum_sin[i] += c[i] * sw;
    }
    v->count++;
    if (v->frame == 0) {
        memcpy(v->mean, c, sizeof(float) * (size_t)nt);
        memset(v->amp, 0, sizeof(float) * (size_t)nt);
        memset(v->phase, 0, sizeof(float) * (size_t)nt);
    }
    if ((iframe || v->count >= v->refit_interval) && v->count > 2) {
        const float inv = 1.0f / v->count;
        for (long i = 0; i < nt; i++) {
            v->mean[i] = v->sum_mean[i] * inv;
            const float sc = v->sum_cos[i] * inv - v->mean[i] * cw;
            const float ss = v->sum_sin[i] * inv - v->mean[i] * sw;
            v->amp[i] = 2.0f * sqrtf(sc * sc + ss * ss);
            v->phase[i] = atan2f(-ss, sc);
        }
        memset(v->sum_cos, 0, sizeof(float) * (size_t)nt);
        memset(v->sum_sin, 0, sizeof(float) * (size_t)nt);
        memset(v->sum_mean, 0, sizeof(float) * (size_t)nt);
        v->count = 0;
    }
    int rc;
    if (iframe) {
        /* n_patches u32, block u8, n_coeffs u16, flags u8 (1: temporal model), origins
         * 3 x i16 per patch, coefficients, omega, mean, amp, phase */
        const uint32_t np = (uint32_t)v->n_patches;
        const uint8_t bs = (uint8_t)v->bs, flags = 1;
        const uint16_t nc = (uint16_t)v->nc;
        const uint64_t fb = (uint64_t)nt * 4;
        const uint64_t len = 8 + (uint64_t)np * 6 + fb + 4 + 3 * fb;
        uint8_t *b = malloc(len);
        if (!b) return -1;
        uint64_t o = 0;
        put(b, &o, &np, 4); put(b, &o, &bs, 1); put(b, &o, &nc, 2); put(b, &o, &flags, 1);
        put(b, &o, v->origins, (uint64_t)np * 6);
        put(b, &o, c, fb);
        put(b, &o, &v->omega, 4);
        put(b, &o, v->mean, fb); put(b, &o, v->amp, fb); put(b, &o, v->phase, fb);
        rc = sfaio_write_vec(out, t, SFAIO_FRAME_VEC_I, b, len);
        free(b);
    } else {
        /* residual against the model; patches whose largest |residual| exceeds the tolerance */
        uint32_t nd = 0;
        for (int p = 0; p < v->n_patches; p++) {
            float mx = 0;
            const long base = (long)p * v->nc;
            float *d = v->delta + (long)nd * v->nc;
            for (int ci = 0; ci < v->nc; ci++) {
                const float pred = v->mean[base + ci] + v->amp[base + ci] * cosf(v->omega * (float)t + v->phase[base + ci]);
                d[ci] = c[base + ci] - pred;
                if (fabsf(d[ci]) > mx) mx = fabsf(d[ci]);
            }
            if (mx > v->delta_tol) v->idx[nd++] = (uint32_t)p;
        }
        /* n_deltas u32, n_coeffs u16, pad u16, indices, residuals */
        const uint16_t nc = (uint16_t)v->nc, pad = 0;
        const uint64_t len = 8 + (uint64_t)nd * 4 + (uint64_t)nd * v->nc * 4;
        uint8_t *b = malloc(len);
        if (!b) return -1;
        uint64_t o = 0;
        put(b, &o, &nd, 4); put(b, &o, &nc, 2); put(b, &o, &pad, 2);
        put(b, &o, v->idx, (uint64_t)nd * 4);
        put(b, &o, v->delta, (uint64_t)nd * v->nc * 4);
        rc = sfaio_write_vec(out, t, SFAIO_FRAME_VEC_P, b, len);
        free(b);
    }
    v->frame++;
    return rc;
}
