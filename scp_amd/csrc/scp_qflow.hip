// scp_qflow.hip — the fixed-charge gradient flow on the GPU (include/scp_amd.h,
// scp_qflow_*): the relaxation of sfa/seed/eta_qflow.c in its fixed-Q mode.
//
// Rotating ansatz Phi_a = phi_a e^{iwt}, Theta_a = theta_a e^{iwt}; at fixed
// charge Q = w N (N = int sum_a |phi_a|^2 + |theta_a|^2) the flow is plain
// gradient descent on E_Q = E_static + Q^2/(2N) with w = Q/N re-derived every
// iteration (eta_qflow.c:242-332):
//   F_phi   = lap phi   - (m^2  - w^2) phi - 2 Vt'(s) phi_a prod_rest + eta curl theta
//   F_theta = lap theta - (mth^2 - w^2) theta + eta curl phi
//   x <- x + dtau F   (Jacobi: every site from the previous iterate)
//
// One launch per iteration updates the 12 fields of every site (7-point
// Laplacian, central curl, periodic wraps) into the other buffer and reduces,
// per block, the new state's norm (the next iteration's N) and max |F| over
// the u and tu forces (the reference's maxF); a one-block pass reduces those
// partials in fixed order into per-iteration device histories, which the
// next launch reads, so a run of many iterations is enqueued without a host
// round trip. At the print iterations an energy pass adds E_Q's terms and
// s_max of the updated state. Every per-site expression keeps the reference's
// order (FP64, no contraction); the reductions run in a fixed order of their
// own, so results agree with the reference to summation-order rounding.
//
// The fixed-(Q, J) mode (eta_qflow.c fixJ = 1) adds, per iteration, a pass
// that forms Jhat w = x d_y w - y d_x w - zhat x w of the 12 fields into a
// second buffer and reduces N, ell = int (u.Jv - v.Ju + tu.Jtv - tv.Jtu) and
// K = int |Jw|^2; a one-thread kernel solves the regularised 2x2 system for the
// two multipliers (omega, Omega; |Omega| capped) into the device history, and
// the step adds the Coriolis and centrifugal forces -2 omega Omega wm (Jhat v)
// - Omega^2 wm Jhat(Jhat u) (v, Theta alike), wm the light-cylinder mask of
// column (i, j), tabulated on the host.
//
// Roofline: HBM; the 12 FP64 fields read and written once per iteration
// (192 B per site-iteration; fixed-(Q, J): 480 B — the fields read twice, the
// Jhat buffer written once and read once), neighbours from L1/L2.
#include "scp_grid.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

using namespace scp_internal;

namespace {

constexpr int QF_BLOCK = 256;
constexpr int QF_NF = 12;   // u0..2, v0..2, tu0..2, tv0..2

struct QfK {
    int N;
    uint32_t n3, fs;
    double invdx2, inv2dx, dV;
    double eta, m, mth, mu, kappa;
    double L, dx;
};

// the fixed-(Q, J) terms of the step: Jhat of the iterate, the mask table, the
// multipliers' histories (null jb: the plain fixed-Q step)
struct QfJ {
    const double *jb, *wtab, *h_omega, *h_Omega;
};

__device__ inline int qwrap(int a, int N) { return a < 0 ? a + N : (a >= N ? a - N : a); }

// deterministic reduction of NV values per thread of a QF_BLOCK block into out[0..NV)
template <int NV>
__device__ inline void qf_block_reduce(double (&v)[NV], const bool (&ismax)[NV], double *out) {
    __shared__ double red[QF_BLOCK / 64][NV];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NV; q++) {
        double x = v[q];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const double y = __shfl_xor(x, o, 64);
            x = ismax[q] ? fmax(x, y) : x + y;
        }
        if (lane == 0) red[wv][q] = x;
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        const int q = threadIdx.x;
        double x = red[0][q];
        for (int w = 1; w < QF_BLOCK / 64; w++) x = ismax[q] ? fmax(x, red[w][q]) : x + red[w][q];
        out[q] = x;
    }
}

// the site's six face neighbours (site indices): X[0][ax] forward, X[1][ax] backward
struct Nb6 {
    uint32_t X[2][3];
};
__device__ inline Nb6 nb6(int N, uint32_t c) {
    const int k = c % N, j = (c / N) % N, i = c / ((uint32_t)N * N);
    const uint32_t NN = (uint32_t)N * N;
    Nb6 b;
    b.X[0][0] = (uint32_t)qwrap(i + 1, N) * NN + j * N + k;
    b.X[1][0] = (uint32_t)qwrap(i - 1, N) * NN + j * N + k;
    b.X[0][1] = i * NN + (uint32_t)qwrap(j + 1, N) * N + k;
    b.X[1][1] = i * NN + (uint32_t)qwrap(j - 1, N) * N + k;
    b.X[0][2] = i * NN + j * N + (uint32_t)qwrap(k + 1, N);
    b.X[1][2] = i * NN + j * N + (uint32_t)qwrap(k - 1, N);
    return b;
}

// iteration `it`: x -> y; hist_n[it] is this iteration's N (set by the previous
// reduction or the initial norm); partials[2 b + {0, 1}]: the new state's raw norm
// sum and max |F| of block b
// (fixed-(Q, J): the multipliers from j.h_omega / j.h_Omega[it], the rotational
// forces from j.jb, the Jhat of x)
template <bool FIXJ>
__global__ void __launch_bounds__(QF_BLOCK)
k_qflow_step(QfK k, const double *__restrict__ x, double *__restrict__ y, double Q, double dtau,
             const double *__restrict__ hist_n, int it, double *__restrict__ partials, QfJ j) {
#pragma clang fp contract(off)
    const uint32_t c = blockIdx.x * QF_BLOCK + threadIdx.x;
    double acc[2] = {0.0, 0.0};
    if (c < k.n3) {
        double omega, Omega = 0.0;
        if (FIXJ) {
            omega = j.h_omega[it];
            Omega = j.h_Omega[it];
        } else {
            omega = Q / hist_n[it];                                    // eta_qflow.c:274
        }
        const double m2w = k.m * k.m - omega * omega, mt2w = k.mth * k.mth - omega * omega;
        const uint32_t fs = k.fs;
        const Nb6 b = nb6(k.N, c);
        const double *u = x, *v = x + 3 * (size_t)fs, *tu = x + 6 * (size_t)fs, *tv = x + 9 * (size_t)fs;
        auto at = [&](const double *f, int comp, uint32_t s) { return f[(size_t)comp * fs + s]; };
        const double u0 = at(u, 0, c), u1 = at(u, 1, c), u2 = at(u, 2, c);
        const double v0 = at(v, 0, c), v1 = at(v, 1, c), v2 = at(v, 2, c);
        const double r0 = u0 * u0 + v0 * v0, r1 = u1 * u1 + v1 * v1, r2 = u2 * u2 + v2 * v2;
        const double s = r0 * r1 * r2, den = 1.0 + k.kappa * s;
        const double Vtp = 0.5 * k.mu / (den * den);
        const double Pa[3] = {r1 * r2, r0 * r2, r0 * r1};
        double xc = 0.0, yc = 0.0, wO = 0.0, O2 = 0.0;
        if (FIXJ) {                                                    // eta_qflow.c:297-300
            const int ci = (int)(c / ((uint32_t)k.N * k.N)), cj = (int)((c / k.N) % k.N);
            xc = -k.L + ci * k.dx;
            yc = -k.L + cj * k.dx;
            const double wm = j.wtab[(uint32_t)ci * k.N + cj];
            wO = 2.0 * omega * Omega * wm;
            O2 = Omega * Omega * wm;
        }
        double nsum = 0.0, maxF = 0.0;
#pragma unroll
        for (int a = 0; a < 3; a++) {
            auto lap = [&](const double *f) {
                return (at(f, a, b.X[0][0]) + at(f, a, b.X[1][0]) + at(f, a, b.X[0][1]) + at(f, a, b.X[1][1]) +
                        at(f, a, b.X[0][2]) + at(f, a, b.X[1][2]) - 6.0 * at(f, a, c)) * k.invdx2;
            };
            const int bb = (a + 1) % 3, cc = (a + 2) % 3;
            auto drv = [&](const double *f, int ax, int comp) { return at(f, comp, b.X[0][ax]) - at(f, comp, b.X[1][ax]); };
            const double lu = lap(u), lv = lap(v), ltu = lap(tu), ltv = lap(tv);
            const double curl_tu = (drv(tu, bb, cc) - drv(tu, cc, bb)) * k.inv2dx;
            const double curl_tv = (drv(tv, bb, cc) - drv(tv, cc, bb)) * k.inv2dx;
            const double curl_u = (drv(u, bb, cc) - drv(u, cc, bb)) * k.inv2dx;
            const double curl_v = (drv(v, bb, cc) - drv(v, cc, bb)) * k.inv2dx;
            const double ua = at(u, a, c), va = at(v, a, c), ta = at(tu, a, c), sa = at(tv, a, c);
            double Fu = lu - m2w * ua - 2.0 * Vtp * ua * Pa[a] + k.eta * curl_tu;
            double Fv = lv - m2w * va - 2.0 * Vtp * va * Pa[a] + k.eta * curl_tv;
            double Ftu = ltu - mt2w * ta + k.eta * curl_u;
            double Ftv = ltv - mt2w * sa + k.eta * curl_v;
            if (FIXJ) {                                                // eta_qflow.c:301-309
                const double *ju = j.jb, *jv = j.jb + 3 * (size_t)fs, *jtu = j.jb + 6 * (size_t)fs,
                             *jtv = j.jb + 9 * (size_t)fs;
                auto jj = [&](const double *w) {     // Jhat of the Jhat field w (component a)
                    const double spin = a == 0 ? at(w, 1, c) : a == 1 ? -at(w, 0, c) : 0.0;
                    return xc * (at(w, a, b.X[0][1]) - at(w, a, b.X[1][1])) * k.inv2dx -
                           yc * (at(w, a, b.X[0][0]) - at(w, a, b.X[1][0])) * k.inv2dx + spin;
                };
                const double j2u = jj(ju), j2v = jj(jv), j2t = jj(jtu), j2s = jj(jtv);
                Fu += -wO * at(jv, a, c) - O2 * j2u;
                Fv += +wO * at(ju, a, c) - O2 * j2v;
                Ftu += -wO * at(jtv, a, c) - O2 * j2t;
                Ftv += +wO * at(jtu, a, c) - O2 * j2s;
            }
            const double nu = ua + dtau * Fu, nv = va + dtau * Fv, ntu = ta + dtau * Ftu, ntv = sa + dtau * Ftv;
            y[(size_t)a * fs + c] = nu;
            y[(size_t)(3 + a) * fs + c] = nv;
            y[(size_t)(6 + a) * fs + c] = ntu;
            y[(size_t)(9 + a) * fs + c] = ntv;
            nsum += nu * nu + nv * nv + ntu * ntu + ntv * ntv;
            maxF = fmax(maxF, fmax(fabs(Fu), fabs(Ftu)));
        }
        acc[0] = nsum;
        acc[1] = maxF;
    }
    const bool ismax[2] = {false, true};
    qf_block_reduce<2>(acc, ismax, partials + 2 * (size_t)blockIdx.x);
}

// Jhat of the 12 fields of x into jb (eta_qflow.c:232-252; spin part zhat x w:
// component 0 gains w_1, component 1 -w_0) and the raw sums into
// partials[3 b + {N, ell, K}]
__global__ void __launch_bounds__(QF_BLOCK)
k_qflow_jhat(QfK k, const double *__restrict__ x, double *__restrict__ jb, double *__restrict__ partials) {
#pragma clang fp contract(off)
    const uint32_t c = blockIdx.x * QF_BLOCK + threadIdx.x;
    double acc[3] = {0.0, 0.0, 0.0};
    if (c < k.n3) {
        const uint32_t fs = k.fs;
        const Nb6 b = nb6(k.N, c);
        const int ci = (int)(c / ((uint32_t)k.N * k.N)), cj = (int)((c / k.N) % k.N);
        const double xc = -k.L + ci * k.dx, yc = -k.L + cj * k.dx;
        auto at = [&](const double *f, int comp, uint32_t s) { return f[(size_t)comp * fs + s]; };
#pragma unroll
        for (int a = 0; a < 3; a++) {
            double jw[4];
#pragma unroll
            for (int q = 0; q < 4; q++) {                              // u, v, tu, tv
                const double *w = x + (size_t)(3 * q) * fs;
                const double spin = a == 0 ? at(w, 1, c) : a == 1 ? -at(w, 0, c) : 0.0;
                jw[q] = xc * (at(w, a, b.X[0][1]) - at(w, a, b.X[1][1])) * k.inv2dx -
                        yc * (at(w, a, b.X[0][0]) - at(w, a, b.X[1][0])) * k.inv2dx + spin;
                jb[(size_t)(3 * q + a) * fs + c] = jw[q];
            }
            const double u = at(x, a, c), v = at(x, 3 + a, c), tu = at(x, 6 + a, c), tv = at(x, 9 + a, c);
            acc[0] += u * u + v * v + tu * tu + tv * tv;
            acc[1] += u * jw[1] - v * jw[0] + tu * jw[3] - tv * jw[2];
            acc[2] += jw[0] * jw[0] + jw[1] * jw[1] + jw[2] * jw[2] + jw[3] * jw[3];
        }
    }
    const bool ismax[3] = {false, false, false};
    qf_block_reduce<3>(acc, ismax, partials + 3 * (size_t)blockIdx.x);
}

// the two multipliers of iteration `it` from its N, ell, K (eta_qflow.c:253-259):
// h_nek[3 it + {0, 1, 2}] -> h_omega[it], h_Omega[it], h_cond[it]
__global__ void k_qflow_solve(const double *__restrict__ h_nek, double *h_omega, double *h_Omega, double *h_cond,
                              double Q, double Jt, int it) {
#pragma clang fp contract(off)
    if (threadIdx.x != 0) return;
    const double OMEGA_CAP = 0.15;
    const double Nn = h_nek[3 * it], ell = h_nek[3 * it + 1], Kk = h_nek[3 * it + 2];
    const double D = Nn * Kk - ell * ell, Dr = D + 1e-10 * Nn * Kk;
    double omega = (Kk * Q - ell * Jt) / Dr;
    double Omega = (ell * Q - Nn * Jt) / Dr;
    if (Omega > OMEGA_CAP) { Omega = OMEGA_CAP; omega = (Q + Omega * ell) / Nn; }
    if (Omega < -OMEGA_CAP) { Omega = -OMEGA_CAP; omega = (Q + Omega * ell) / Nn; }
    h_omega[it] = omega;
    h_Omega[it] = Omega;
    h_cond[it] = D / (Nn * Kk);
}

// raw norm sums of a state into partials[3 b + {total, phi, theta}]; the total
// groups each (site, a) as the flow's N does (eta_qflow.c:272-273)
__global__ void __launch_bounds__(QF_BLOCK) k_qflow_norm(QfK k, const double *__restrict__ x, double *partials) {
#pragma clang fp contract(off)
    const uint32_t c = blockIdx.x * QF_BLOCK + threadIdx.x;
    double acc[3] = {0.0, 0.0, 0.0};
    if (c < k.n3) {
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const double u = x[(size_t)a * k.fs + c], v = x[(size_t)(3 + a) * k.fs + c];
            const double tu = x[(size_t)(6 + a) * k.fs + c], tv = x[(size_t)(9 + a) * k.fs + c];
            acc[0] += u * u + v * v + tu * tu + tv * tv;
            acc[1] += u * u + v * v;
            acc[2] += tu * tu + tv * tv;
        }
    }
    const bool ismax[3] = {false, false, false};
    qf_block_reduce<3>(acc, ismax, partials + 3 * (size_t)blockIdx.x);
}

// E_Q terms of a state (eta_qflow.c:338-368): partials[5 b + {Eg, Em, Ep, Ex, s_max}]
__global__ void __launch_bounds__(QF_BLOCK)
k_qflow_energy(QfK k, const double *__restrict__ x, double *__restrict__ partials) {
#pragma clang fp contract(off)
    const uint32_t c = blockIdx.x * QF_BLOCK + threadIdx.x;
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (c < k.n3) {
        const uint32_t fs = k.fs;
        const Nb6 b = nb6(k.N, c);
        const double *u = x, *v = x + 3 * (size_t)fs, *tu = x + 6 * (size_t)fs, *tv = x + 9 * (size_t)fs;
        auto at = [&](const double *f, int comp, uint32_t s) { return f[(size_t)comp * fs + s]; };
        const double r0 = at(u, 0, c) * at(u, 0, c) + at(v, 0, c) * at(v, 0, c);
        const double r1 = at(u, 1, c) * at(u, 1, c) + at(v, 1, c) * at(v, 1, c);
        const double r2 = at(u, 2, c) * at(u, 2, c) + at(v, 2, c) * at(v, 2, c);
        const double s = r0 * r1 * r2;
        acc[4] = s;
        acc[2] = 0.5 * k.mu * s / (1.0 + k.kappa * s);
        double Eg = 0.0, Em = 0.0, Ex = 0.0;
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const double *F4[4] = {u, v, tu, tv};
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const double *f = F4[q];
                const double gx = (at(f, a, b.X[0][0]) - at(f, a, b.X[1][0])) * k.inv2dx;
                const double gy = (at(f, a, b.X[0][1]) - at(f, a, b.X[1][1])) * k.inv2dx;
                const double gz = (at(f, a, b.X[0][2]) - at(f, a, b.X[1][2])) * k.inv2dx;
                Eg += 0.5 * (gx * gx + gy * gy + gz * gz);
            }
            const double ua = at(u, a, c), va = at(v, a, c), ta = at(tu, a, c), sa = at(tv, a, c);
            Em += 0.5 * k.m * k.m * (ua * ua + va * va) + 0.5 * k.mth * k.mth * (ta * ta + sa * sa);
            const int bb = (a + 1) % 3, cc = (a + 2) % 3;
            auto drv = [&](const double *f, int ax, int comp) { return at(f, comp, b.X[0][ax]) - at(f, comp, b.X[1][ax]); };
            const double curl_u = (drv(u, bb, cc) - drv(u, cc, bb)) * k.inv2dx;
            const double curl_v = (drv(v, bb, cc) - drv(v, cc, bb)) * k.inv2dx;
            Ex += -k.eta * (ta * curl_u + sa * curl_v);
        }
        acc[0] = Eg;
        acc[1] = Em;
        acc[3] = Ex;
    }
    const bool ismax[5] = {false, false, false, false, true};
    qf_block_reduce<5>(acc, ismax, partials + 5 * (size_t)blockIdx.x);
}

// destinations of the reduced slots (null: not kept)
struct QfOut {
    double *p[5];
};

// one block: rows b of width NV in fixed order (thread t: rows t, t + 256, ...,
// then the block reduction); slot q goes to o.p[q], sums times sumscale
template <int NV>
__global__ void __launch_bounds__(QF_BLOCK)
k_qflow_reduce(const double *__restrict__ partials, int nblocks, int maxmask, double sumscale, QfOut o) {
#pragma clang fp contract(off)
    double acc[NV];
    bool ismax[NV];
#pragma unroll
    for (int q = 0; q < NV; q++) {
        acc[q] = 0.0;
        ismax[q] = (maxmask >> q) & 1;
    }
    for (int b = threadIdx.x; b < nblocks; b += QF_BLOCK)
#pragma unroll
        for (int q = 0; q < NV; q++) {
            const double y = partials[(size_t)b * NV + q];
            acc[q] = ismax[q] ? fmax(acc[q], y) : acc[q] + y;
        }
    __shared__ double out[NV];
    qf_block_reduce<NV>(acc, ismax, out);
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NV; q++)
            if (o.p[q]) *o.p[q] = ismax[q] ? out[q] : out[q] * sumscale;
    }
}

}  // namespace

struct scp_qflow {
    scp_qflow_params p;
    QfK k;
    hipStream_t stream = nullptr;
    double *x[2] = {nullptr, nullptr};
    int cur = 0;
    int nblocks = 0;
    double *partials = nullptr;   // max(2, 3, 5) x nblocks
    double *scal = nullptr;       // 3: norms (total, phi, theta)
    double *jb = nullptr;         // fixed-(Q, J): Jhat of the 12 fields (12 x fs), from the first such run
    double *wtab = nullptr;       // fixed-(Q, J): the light-cylinder mask per column (N x N)
    bool jb_valid = false;        // jb holds Jhat of the current state
};

static int qf_err(hipError_t e, const char *what) {
    return e == hipSuccess ? 0 : set_err(-2, "%s failed: %s", what, hipGetErrorString(e));
}

static void qf_free(scp_qflow *f) {
    if (!f) return;
    hipFree(f->x[0]); hipFree(f->x[1]); hipFree(f->partials); hipFree(f->scal);
    hipFree(f->jb); hipFree(f->wtab);
    if (f->stream) hipStreamDestroy(f->stream);
    delete f;
}

extern "C" int scp_qflow_create(const scp_qflow_params *p, scp_qflow **out) {
    if (!p || !out) return set_err(-1, "null argument");
    *out = nullptr;
    if (p->N < 4 || p->N > 1024) return set_err(-1, "N=%d out of range [4, 1024]", p->N);
    if (!(p->L > 0)) return set_err(-1, "L must be positive");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return set_err(-4, "no HIP device");
    scp_qflow *f = new scp_qflow;
    f->p = *p;
    const int N = p->N;
    const double dx = 2.0 * p->L / (N - 1), dx2 = dx * dx;       // eta_qflow.c:102
    f->k.N = N;
    f->k.n3 = (uint32_t)((int64_t)N * N * N);
    f->k.fs = (uint32_t)(((int64_t)f->k.n3 + 63) / 64 * 64);
    f->k.invdx2 = 1.0 / dx2;
    f->k.inv2dx = 1.0 / (2.0 * dx);
    f->k.dV = dx * dx * dx;
    f->k.eta = p->eta; f->k.m = p->m; f->k.mth = p->mth; f->k.mu = p->mu; f->k.kappa = p->kappa;
    f->k.L = p->L;
    f->k.dx = dx;
    f->nblocks = (int)((f->k.n3 + QF_BLOCK - 1) / QF_BLOCK);
    const size_t bytes = (size_t)QF_NF * f->k.fs * sizeof(double);
    hipError_t e = hipStreamCreateWithFlags(&f->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(&f->x[0], bytes);
    if (e == hipSuccess) e = hipMalloc(&f->x[1], bytes);
    if (e == hipSuccess) e = hipMalloc(&f->partials, (size_t)5 * f->nblocks * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&f->scal, 3 * sizeof(double));
    if (e != hipSuccess) {
        qf_free(f);
        return set_err(-2, "qflow allocation failed: %s", hipGetErrorString(e));
    }
    *out = f;
    return 0;
}

extern "C" int scp_qflow_destroy(scp_qflow *f) {
    if (!f) return 0;
    hipStreamSynchronize(f->stream);
    qf_free(f);
    return 0;
}

extern "C" int scp_qflow_upload(scp_qflow *f, const double *const *cols, int ncols) {
    if (!f || !cols) return set_err(-1, "null argument");
    if (ncols != QF_NF) return set_err(-1, "ncols=%d, the flow holds 12 fields (u, v, tu, tv x 3)", ncols);
    const size_t nb = (size_t)f->k.n3 * sizeof(double);
    for (int q = 0; q < QF_NF; q++) {
        if (!cols[q]) return set_err(-1, "null column %d", q);
        int rc = qf_err(hipMemcpyAsync(f->x[0] + (size_t)q * f->k.fs, cols[q], nb, hipMemcpyHostToDevice, f->stream),
                        "upload");
        if (rc) return rc;
    }
    f->cur = 0;
    f->jb_valid = false;
    return qf_err(hipStreamSynchronize(f->stream), "upload sync");
}

extern "C" int scp_qflow_download(scp_qflow *f, double *const *cols, int ncols) {
    if (!f || !cols) return set_err(-1, "null argument");
    if (ncols != QF_NF) return set_err(-1, "ncols=%d, the flow holds 12 fields (u, v, tu, tv x 3)", ncols);
    const size_t nb = (size_t)f->k.n3 * sizeof(double);
    for (int q = 0; q < QF_NF; q++) {
        if (!cols[q]) return set_err(-1, "null column %d", q);
        int rc = qf_err(hipMemcpyAsync(cols[q], f->x[f->cur] + (size_t)q * f->k.fs, nb, hipMemcpyDeviceToHost,
                                       f->stream), "download");
        if (rc) return rc;
    }
    return qf_err(hipStreamSynchronize(f->stream), "download sync");
}

// norms of the current state into dst[0..3) (total, phi, theta; each times dV)
static int enqueue_norm(scp_qflow *f, double *d_total, double *d_phi, double *d_theta) {
    hipLaunchKernelGGL(k_qflow_norm, dim3(f->nblocks), dim3(QF_BLOCK), 0, f->stream, f->k, f->x[f->cur], f->partials);
    QfOut o{{d_total, d_phi, d_theta, nullptr, nullptr}};
    hipLaunchKernelGGL(k_qflow_reduce<3>, dim3(1), dim3(QF_BLOCK), 0, f->stream, f->partials, f->nblocks, 0, f->k.dV, o);
    return qf_err(hipGetLastError(), "norm launch");
}

extern "C" int scp_qflow_norm(scp_qflow *f, double *n_phi, double *n_theta) {
    if (!f) return set_err(-1, "null flow");
    int rc = enqueue_norm(f, f->scal, f->scal + 1, f->scal + 2);
    if (rc) return rc;
    double h[3];
    rc = qf_err(hipMemcpyAsync(h, f->scal, sizeof h, hipMemcpyDeviceToHost, f->stream), "norm copy");
    if (!rc) rc = qf_err(hipStreamSynchronize(f->stream), "norm sync");
    if (rc) return rc;
    if (n_phi) *n_phi = h[1];
    if (n_theta) *n_theta = h[2];
    return 0;
}

extern "C" int scp_qflow_run(scp_qflow *f, double Q, double dtau, int iters, int printev, scp_qflow_row *rows,
                             int maxrows, int *nrows, double *omega_last) {
    if (!f) return set_err(-1, "null flow");
    if (iters < 0) return set_err(-1, "iters=%d must be >= 0", iters);
    if (printev < 1) return set_err(-1, "printev=%d must be >= 1", printev);
    if (!(Q > 0) || !(dtau > 0)) return set_err(-1, "Q and dtau must be positive");
    if (maxrows < 0 || (maxrows > 0 && !rows)) return set_err(-1, "bad row buffer");
    const int nit = iters + 1;                                  // it = 0 .. iters (eta_qflow.c:242)
    std::vector<int> pits;
    for (int it = 0; it <= iters; it++)
        if (it % printev == 0 || it == iters) pits.push_back(it);
    // device histories: N entering each iteration (and after the last), max|F| of
    // each iteration, the energy terms {Eg, Em, Ep, Ex, s_max} of each print row
    const size_t nh = (size_t)(nit + 1) + nit + 5 * pits.size();
    double *hist = nullptr;
    int rc = qf_err(hipMalloc(&hist, nh * sizeof(double)), "qflow history alloc");
    if (rc) return rc;
    double *hN = hist, *hF = hist + nit + 1, *hE = hF + nit;
    rc = enqueue_norm(f, hN, nullptr, nullptr);
    size_t pr = 0;
    for (int it = 0; it < nit && !rc; it++) {
        hipLaunchKernelGGL(k_qflow_step<false>, dim3(f->nblocks), dim3(QF_BLOCK), 0, f->stream, f->k,
                           f->x[f->cur], f->x[f->cur ^ 1], Q, dtau, hN, it, f->partials, QfJ{});
        QfOut o{{hN + it + 1, hF + it, nullptr, nullptr, nullptr}};
        hipLaunchKernelGGL(k_qflow_reduce<2>, dim3(1), dim3(QF_BLOCK), 0, f->stream, f->partials, f->nblocks, 2,
                           f->k.dV, o);
        f->cur ^= 1;
        if (pr < pits.size() && pits[pr] == it) {              // E_Q of the updated state (eta_qflow.c:336-373)
            hipLaunchKernelGGL(k_qflow_energy, dim3(f->nblocks), dim3(QF_BLOCK), 0, f->stream, f->k, f->x[f->cur],
                               f->partials);
            double *e = hE + 5 * pr;
            QfOut oe{{e, e + 1, e + 2, e + 3, e + 4}};
            hipLaunchKernelGGL(k_qflow_reduce<5>, dim3(1), dim3(QF_BLOCK), 0, f->stream, f->partials, f->nblocks,
                               16, 1.0, oe);
            pr++;
        }
        rc = qf_err(hipGetLastError(), "qflow launch");
    }
    std::vector<double> h(nh);
    if (!rc) rc = qf_err(hipMemcpyAsync(h.data(), hist, nh * sizeof(double), hipMemcpyDeviceToHost, f->stream),
                         "history copy");
    if (!rc) rc = qf_err(hipStreamSynchronize(f->stream), "qflow sync");
    hipFree(hist);
    if (rc) return rc;
    const double dV = f->k.dV;
    const double *N_ = h.data(), *F_ = N_ + nit + 1, *E_ = F_ + nit;
    int nr = 0;
    for (size_t r = 0; r < pits.size(); r++) {
        const int it = pits[r];
        const double Nn = N_[it], *e = E_ + 5 * r;
        if (nr < maxrows) {
            scp_qflow_row &w = rows[nr];
            w.iter = it;
            w.reserved = 0;
            w.omega = Q / Nn;
            w.E_Q = (e[0] + e[1] + e[2] + e[3]) * dV + Q * Q / (2.0 * Nn);   // eta_qflow.c:369
            w.maxF = F_[it];
            w.s_max = e[4];
            w.N = Nn;
        }
        nr++;
    }
    if (nrows) *nrows = nr;
    if (omega_last) *omega_last = Q / N_[iters];
    f->jb_valid = false;
    return 0;
}

extern "C" int scp_qflow_run_fixj(scp_qflow *f, double Q, double Jt, double dtau, int iters, int printev,
                                  scp_qflow_row_j *rows, int maxrows, int *nrows, double *last) {
    if (!f) return set_err(-1, "null flow");
    if (iters < 0) return set_err(-1, "iters=%d must be >= 0", iters);
    if (printev < 1) return set_err(-1, "printev=%d must be >= 1", printev);
    if (!(Q > 0) || !(dtau > 0)) return set_err(-1, "Q and dtau must be positive");
    if (!std::isfinite(Jt)) return set_err(-1, "J target must be finite");
    if (maxrows < 0 || (maxrows > 0 && !rows)) return set_err(-1, "bad row buffer");
    const int N = f->k.N;
    int rc = 0;
    if (!f->jb) {
        rc = qf_err(hipMalloc(&f->jb, (size_t)QF_NF * f->k.fs * sizeof(double)), "qflow Jhat alloc");
        if (!rc) rc = qf_err(hipMalloc(&f->wtab, (size_t)N * N * sizeof(double)), "qflow mask alloc");
        if (rc) return rc;
        std::vector<double> w((size_t)N * N);                  // eta_qflow.c:213-217
        for (int i = 0; i < N; i++)
            for (int j = 0; j < N; j++) {
                const double x = -f->p.L + i * f->k.dx, y = -f->p.L + j * f->k.dx, rho = std::sqrt(x * x + y * y);
                w[(size_t)i * N + j] = 1.0 / (1.0 + std::exp((rho - 11.0) / 0.5));
            }
        rc = qf_err(hipMemcpyAsync(f->wtab, w.data(), w.size() * sizeof(double), hipMemcpyHostToDevice, f->stream),
                    "qflow mask upload");
        if (!rc) rc = qf_err(hipStreamSynchronize(f->stream), "qflow mask sync");
        if (rc) return rc;
    }
    const int nit = iters + 1;
    std::vector<int> pits;
    for (int it = 0; it <= iters; it++)
        if (it % printev == 0 || it == iters) pits.push_back(it);
    // device histories: {N, ell, K} of each iteration's iterate and of the final
    // state, omega / Omega / cond and max|F| of each iteration, the energy terms
    // {Eg, Em, Ep, Ex, s_max} of each print row
    const size_t nh = (size_t)3 * (nit + 1) + 4 * (size_t)nit + 5 * pits.size();
    double *hist = nullptr;
    rc = qf_err(hipMalloc(&hist, nh * sizeof(double)), "qflow history alloc");
    if (rc) return rc;
    double *hNEK = hist, *hOm = hNEK + 3 * (size_t)(nit + 1), *hOO = hOm + nit, *hC = hOO + nit, *hF = hC + nit;
    double *hE = hF + nit;
    size_t pr = 0;
    for (int it = 0; it < nit && !rc; it++) {
        double *nek = hNEK + 3 * (size_t)it;
        hipLaunchKernelGGL(k_qflow_jhat, dim3(f->nblocks), dim3(QF_BLOCK), 0, f->stream, f->k, f->x[f->cur], f->jb,
                           f->partials);
        QfOut oj{{nek, nek + 1, nek + 2, nullptr, nullptr}};
        hipLaunchKernelGGL(k_qflow_reduce<3>, dim3(1), dim3(QF_BLOCK), 0, f->stream, f->partials, f->nblocks, 0,
                           f->k.dV, oj);
        hipLaunchKernelGGL(k_qflow_solve, dim3(1), dim3(64), 0, f->stream, hNEK, hOm, hOO, hC, Q, Jt, it);
        QfJ jj{f->jb, f->wtab, hOm, hOO};
        hipLaunchKernelGGL(k_qflow_step<true>, dim3(f->nblocks), dim3(QF_BLOCK), 0, f->stream, f->k, f->x[f->cur],
                           f->x[f->cur ^ 1], Q, dtau, nullptr, it, f->partials, jj);
        QfOut o{{nullptr, hF + it, nullptr, nullptr, nullptr}};
        hipLaunchKernelGGL(k_qflow_reduce<2>, dim3(1), dim3(QF_BLOCK), 0, f->stream, f->partials, f->nblocks, 2,
                           f->k.dV, o);
        f->cur ^= 1;
        if (pr < pits.size() && pits[pr] == it) {              // E_QJ of the updated state (eta_qflow.c:336-358)
            hipLaunchKernelGGL(k_qflow_energy, dim3(f->nblocks), dim3(QF_BLOCK), 0, f->stream, f->k, f->x[f->cur],
                               f->partials);
            double *e = hE + 5 * pr;
            QfOut oe{{e, e + 1, e + 2, e + 3, e + 4}};
            hipLaunchKernelGGL(k_qflow_reduce<5>, dim3(1), dim3(QF_BLOCK), 0, f->stream, f->partials, f->nblocks,
                               16, 1.0, oe);
            pr++;
        }
        rc = qf_err(hipGetLastError(), "qflow launch");
    }
    if (!rc) {                                                 // Jhat of the final state (eta_qflow.c:377-399)
        double *nek = hNEK + 3 * (size_t)nit;
        hipLaunchKernelGGL(k_qflow_jhat, dim3(f->nblocks), dim3(QF_BLOCK), 0, f->stream, f->k, f->x[f->cur], f->jb,
                           f->partials);
        QfOut oj{{nek, nek + 1, nek + 2, nullptr, nullptr}};
        hipLaunchKernelGGL(k_qflow_reduce<3>, dim3(1), dim3(QF_BLOCK), 0, f->stream, f->partials, f->nblocks, 0,
                           f->k.dV, oj);
        rc = qf_err(hipGetLastError(), "qflow launch");
    }
    std::vector<double> h(nh);
    if (!rc) rc = qf_err(hipMemcpyAsync(h.data(), hist, nh * sizeof(double), hipMemcpyDeviceToHost, f->stream),
                         "history copy");
    if (!rc) rc = qf_err(hipStreamSynchronize(f->stream), "qflow sync");
    hipFree(hist);
    if (rc) return rc;
    const double dV = f->k.dV;
    const double *NEK = h.data(), *Om_ = NEK + 3 * (size_t)(nit + 1), *OO_ = Om_ + nit, *C_ = OO_ + nit;
    const double *F_ = C_ + nit, *E_ = F_ + nit;
    int nr = 0;
    for (size_t r = 0; r < pits.size(); r++) {
        const int it = pits[r];
        const double *nek = NEK + 3 * (size_t)it, *e = E_ + 5 * r;
        if (nr < maxrows) {
            scp_qflow_row_j &w = rows[nr];
            w.iter = it;
            w.reserved = 0;
            w.omega = Om_[it];
            w.Omega = OO_[it];
            w.E_QJ = (e[0] + e[1] + e[2] + e[3]) * dV + 0.5 * (Om_[it] * Q - OO_[it] * Jt);   // eta_qflow.c:354
            w.maxF = F_[it];
            w.s_max = e[4];
            w.N = nek[0];
            w.J_act = Om_[it] * nek[1] - OO_[it] * nek[2];                                  // eta_qflow.c:355
            w.cond = C_[it];
        }
        nr++;
    }
    if (nrows) *nrows = nr;
    if (last) {
        last[0] = Om_[iters];
        last[1] = OO_[iters];
        last[2] = C_[iters];
        last[3] = NEK[3 * (size_t)nit + 1];
        last[4] = NEK[3 * (size_t)nit + 2];
    }
    f->jb_valid = true;
    return 0;
}

extern "C" int scp_qflow_download_jhat(scp_qflow *f, double *const *cols, int ncols) {
    if (!f || !cols) return set_err(-1, "null argument");
    if (ncols != QF_NF) return set_err(-1, "ncols=%d, Jhat holds 12 fields (u, v, tu, tv x 3)", ncols);
    if (!f->jb_valid) return set_err(-1, "no Jhat of the current state: scp_qflow_run_fixj has not run on it");
    const size_t nb = (size_t)f->k.n3 * sizeof(double);
    for (int q = 0; q < QF_NF; q++) {
        if (!cols[q]) return set_err(-1, "null column %d", q);
        int rc = qf_err(hipMemcpyAsync(cols[q], f->jb + (size_t)q * f->k.fs, nb, hipMemcpyDeviceToHost, f->stream),
                        "Jhat download");
        if (rc) return rc;
    }
    return qf_err(hipStreamSynchronize(f->stream), "Jhat download sync");
}
