// scp_comp32.h — compensated FP32 arithmetic of the FP32C step and the
// leapfrog update of one scalar, shared by the LDS-staged march
// (scp_march.h) and the direct-neighbour kernel (scp_step32.hip).
//
// A scalar is a pair (hi, lo) of FP32 words, |lo| <= ulp(hi)/2. Increments
// are formed with FMAs from the FP64 step constants split into FP32 pairs
// and added by Knuth's TwoSum, which is exact for any magnitudes (velocities
// and fields pass through zero, so |hi| >= |increment| cannot be assumed).
// The kernels that include this are compiled with FMA contraction allowed:
// the sums contain no products (increments use explicit fmaf), so
// contraction cannot change them.
#pragma once

#include "scp_kernels.h"

#include <hip/hip_runtime.h>

namespace scp {

constexpr float TWO_PI_HI = 6.28318548202514648f;      // fl32(2 pi)
constexpr float TWO_PI_LO = -1.7484555314695172e-07f;  // 2 pi - TWO_PI_HI
constexpr float INV_TWO_PI = 0.159154943091895336f;

__device__ inline void fast_two_sum(float a, float b, float &s, float &e) {
    s = a + b;
    e = b - (s - a);
}

// A constant as an unevaluated pair c = h + l of FP32 words (the FP64 step
// constants dt, dt/2, -g dx dt split once), so increments are formed with the
// FP64 constant and not with its FP32 rounding.
struct DF {
    float h, l;
};
__device__ inline DF df_split(double c) {
    const float h = (float)c;
    return {h, (float)(c - (double)h)};
}

// (hi, lo) += c * (xh + xl)
__device__ inline void comp_add(float &hi, float &lo, DF c, float xh, float xl) {
    const float y = fmaf(c.h, xh, fmaf(c.l, xh, fmaf(c.h, xl, lo)));
    const float t = hi + y;
    const float bp = t - hi;
    lo = (hi - (t - bp)) + (y - bp);
    hi = t;
}
// (hi, lo) += c * a for a plain FP32 a (a kick by an acceleration)
__device__ inline void comp_kick(float &hi, float &lo, DF c, float a) {
    const float y = fmaf(c.h, a, fmaf(c.l, a, lo));
    const float t = hi + y;
    const float bp = t - hi;
    lo = (hi - (t - bp)) + (y - bp);
    hi = t;
}
// (hi, lo) *= d
__device__ inline void df_scale(float &hi, float &lo, float d) {
    const float p = __fmul_rn(hi, d);
    const float e = fmaf(hi, d, -p) + __fmul_rn(lo, d);
    fast_two_sum(p, e, hi, lo);
}

// One scalar's step (the reference's stages, scp_sim.c:1583-1716): from the
// stored W = v_n + dt/2 a_n and x_{n+1} with a = a_{n+1},
//   v  = W + dt/2 a            completed velocity (stage 4), times the sponge
//                              factor d where damped (stage 5);
//   W' = v + dt/2 a            next first half kick (stage 1);
//   x' = x + dt W'             next drift (stage 2), or for links
//                              x' = wrap(x + gad W') into [-pi, pi].
// Undamped, the two half kicks are one compensated add W' = W + dt a
// (dt = 2 (dt/2) exactly, so the FP32 pair of dt is twice that of dt/2), and v
// is formed only for callers that read it (emitted state, diagnostics);
// either way W' and x' do not depend on whether v is read, so a run gives the
// same state however it is split into calls.
__device__ inline void leap32(bool link, float a, bool damp, float d, const DF &hdt, const DF &dt, const DF &gad,
                              float &wh, float &wl, float &xh, float &xl, float &vh, float &vl) {
    vh = wh; vl = wl;
    comp_kick(vh, vl, hdt, a);
    if (damp) {
        df_scale(vh, vl, d);
        wh = vh; wl = vl;
        comp_kick(wh, wl, hdt, a);
    } else {
        comp_kick(wh, wl, dt, a);
    }
    if (link) {
        comp_add(xh, xl, gad, wh, wl);
        const float n = rintf(xh * INV_TWO_PI);
        if (n != 0.0f) comp_add(xh, xl, DF{-n, 0.0f}, TWO_PI_HI, TWO_PI_LO);
    } else {
        comp_add(xh, xl, dt, wh, wl);
    }
}

}  // namespace scp
