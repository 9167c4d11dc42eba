// scp_seed.hip — Gaussian Q-ball initial states generated on the device
// (scp_grid_seed_balls), the device form of the host generator in
// csrc/host/seeds.c: the same envelope, phases, cut-off box and
// counter-based noise, so a job described by (config, balls, seed) needs no
// host-side state and no host-to-device copy of N^3 columns.
#include "scp_grid.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

namespace scp {

namespace {

constexpr double TWO_PI_D = 6.283185307179586476925286766559;
constexpr double CUTOFF_SIGMAS = 10.0;   // exp(-50) ~ 2e-22: contributions beyond are dropped

__device__ inline uint64_t mix64(uint64_t z) {   // splitmix64 finaliser
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__device__ inline double rng_u01(uint64_t seed, uint64_t stream, uint64_t index) {
    uint64_t h = mix64(mix64(seed) ^ (stream * 0xD1B54A32D192ED03ull));
    h = mix64(h ^ index);
    return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}
__device__ inline double rng_normal(uint64_t seed, uint64_t stream, uint64_t index) {   // Box-Muller
    const double u1 = 1.0 - rng_u01(seed, 2 * stream, index);
    const double u2 = rng_u01(seed, 2 * stream + 1, index);
    return sqrt(-2.0 * log(u1)) * cos(TWO_PI_D * u2);
}
__device__ inline double min_image(double d, double box) {
    d = fmod(d, box);
    if (d > 0.5 * box) d -= box;
    else if (d < -0.5 * box) d += box;
    return d;
}

// per-ball constants prepared on the host
struct BallDev {
    double x0, y0, z0, amp, inv2s2, omega, cre[3], cim[3];
    int32_t ci, cj, ck, half;   // centre indices and cut-off half-width (-1: whole grid)
};

__global__ void k_seed(int N, double L, double dx, uint32_t fs, int gauged, const BallDev *balls, int nballs,
                       double theta_sigma, double link_sigma, uint64_t seed, double *xo, double *vo) {
    const uint32_t n3 = (uint32_t)N * N * N;
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n3) return;
    const int k = c % N, j = (c / N) % N, i = c / ((uint32_t)N * N);
    const double box = N * dx;
    const double x = -L + i * dx, y = -L + j * dx, z = -L + k * dx;
    double u[3] = {0, 0, 0}, v[3] = {0, 0, 0}, ud[3] = {0, 0, 0}, vd[3] = {0, 0, 0};
    auto inside = [&](int idx, int cidx, int half) {
        const int off = ((idx - cidx) % N + N) % N;
        return off <= half || off >= N - half;
    };
    for (int b = 0; b < nballs; b++) {       // ball order as on the host: sums round identically
        const BallDev B = balls[b];
        if (B.half >= 0 && !(inside(i, B.ci, B.half) && inside(j, B.cj, B.half) && inside(k, B.ck, B.half)))
            continue;
        const double ex = min_image(x - B.x0, box), ey = min_image(y - B.y0, box), ez = min_image(z - B.z0, box);
        const double r2 = ex * ex + ey * ey + ez * ez;
        const double env = B.amp * exp(-r2 * B.inv2s2);
        for (int a = 0; a < 3; a++) {
            const double re = env * B.cre[a], im = env * B.cim[a];
            u[a] += re;
            v[a] += im;
            ud[a] += -B.omega * im;   // d/dt (re + i im) = i omega (re + i im)
            vd[a] += B.omega * re;
        }
    }
    for (int a = 0; a < 3; a++) {
        xo[(size_t)(F_U + a) * fs + c] = u[a];
        xo[(size_t)(F_V + a) * fs + c] = v[a];
        xo[(size_t)(F_TU + a) * fs + c] = theta_sigma * rng_normal(seed, (uint64_t)a, c);
        xo[(size_t)(F_TV + a) * fs + c] = theta_sigma * rng_normal(seed, (uint64_t)(3 + a), c);
        vo[(size_t)(F_U + a) * fs + c] = ud[a];
        vo[(size_t)(F_V + a) * fs + c] = vd[a];
        vo[(size_t)(F_TU + a) * fs + c] = 0.0;
        vo[(size_t)(F_TV + a) * fs + c] = 0.0;
    }
    if (gauged) {
        for (int d = 0; d < 3; d++) {
            xo[(size_t)(F_TH + d) * fs + c] = link_sigma == 0.0 ? 0.0 : link_sigma * rng_normal(seed, (uint64_t)(6 + d), c);
            vo[(size_t)(F_TH + d) * fs + c] = 0.0;
        }
    }
}

}  // namespace

// host side: ball constants, the launch, and the device copy of the ball list
int seed_balls_device(int N, double L, double dx, uint32_t fs, int gauged, const scp_qball *balls, int nballs,
                      double theta_sigma, double link_sigma, uint64_t seed, double *xo, double *vo,
                      hipStream_t s, const scp_internal::GridMem *mem) {
    std::vector<BallDev> bd((size_t)(nballs > 0 ? nballs : 1));
    for (int b = 0; b < nballs; b++) {
        const scp_qball &B = balls[b];
        BallDev &D = bd[(size_t)b];
        D.x0 = B.x0; D.y0 = B.y0; D.z0 = B.z0; D.amp = B.amp; D.omega = B.omega;
        D.inv2s2 = 1.0 / (2.0 * B.sigma * B.sigma);
        for (int a = 0; a < 3; a++) {
            const double ph = B.phase + B.winding * TWO_PI_D * a / 3.0;
            D.cre[a] = cos(ph);
            D.cim[a] = sin(ph);
        }
        const int half = (int)ceil(CUTOFF_SIGMAS * B.sigma / dx);
        const bool whole = 2 * half + 1 >= N;
        D.half = whole ? -1 : half;
        D.ci = whole ? 0 : (int)lround((B.x0 + L) / dx);
        D.cj = whole ? 0 : (int)lround((B.y0 + L) / dx);
        D.ck = whole ? 0 : (int)lround((B.z0 + L) / dx);
    }
    BallDev *d = nullptr;
    hipError_t e = mem ? mem->alloc(mem->ctx, (void **)&d, bd.size() * sizeof(BallDev))
                       : hipMalloc((void **)&d, bd.size() * sizeof(BallDev));
    if (e != hipSuccess) return (int)e;
    e = hipMemcpyAsync(d, bd.data(), bd.size() * sizeof(BallDev), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) {
        const uint32_t n3 = (uint32_t)N * N * N;
        hipLaunchKernelGGL(k_seed, dim3((n3 + 255) / 256), dim3(256), 0, s, N, L, dx, fs, gauged, d, nballs,
                           theta_sigma, link_sigma, seed, xo, vo);
        e = hipGetLastError();
    }
    const hipError_t e2 = hipStreamSynchronize(s);   // the ball list is read by this launch only
    if (mem) mem->release(mem->ctx, d);
    else hipFree(d);
    return (int)(e != hipSuccess ? e : e2);
}

}  // namespace scp
