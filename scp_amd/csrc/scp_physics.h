// scp_physics.h — per-site physics of the sfa/sim step, shared by the FP64
// kernels (R = double, compiled with -ffp-contract=off: every expression tree
// is the reference's, so results round identically) and the FP32 kernels
// (R = float). Geometry helpers of the z-march decomposition live here too.
#pragma once

#include "scp_kernels.h"

#include <hip/hip_runtime.h>

#include <type_traits>

namespace scp {

// FP32 stencils use difference forms: sum_n (x_n - x) instead of sum_n x_n -
// 6x, and transported differences U x_n - x = (x_n - x) + (U - 1) x_n with
// cos(th) - 1 formed without cancellation. For smooth fields x_n - x is exact
// in FP32 (Sterbenz), so the Laplacian carries a rounding error relative to
// itself instead of to the field (~|x| eps h2 per site). FP64 keeps the
// reference's expression trees (and so its rounding) exactly.
template <typename R>
constexpr bool diff_form() { return std::is_same<R, float>::value; }

#define SCP_PI 3.14159265358979323846

// FP32 sine, cosine and cos - 1 of a link or plaquette angle. x = k pi/2 + r
// with pi/2 split into three FP32 parts (Cody-Waite, each step one FMA), so
// |r| <= pi/4 carries about one rounding of r whatever k is; sin r and cos r
// are their Taylor series to r^9 and r^10 (truncation below 2e-9 and 2e-10 of
// the value on |r| <= pi/4, under half an FP32 ulp) by Horner's rule with FMAs;
// k mod 4 selects and signs them. cos x - 1 keeps its relative accuracy near
// x = 0: for k = 0 it is the cosine series without its constant term, and in
// the other quadrants cos x - 1 is below -0.29, where the subtraction is exact
// to one rounding. Measured against double precision: sin and cos within 1.5
// ulp, cos - 1 within 3.5 ulp for |x| <= 100 (link angles lie in [-pi, pi],
// plaquette angles in [-4 pi, 4 pi]).
struct SinCos32 { float s, c, cm1; };
__device__ inline SinCos32 sincos32(float x) {
    constexpr float TWO_OVER_PI = 0.636619746685028076f;
    constexpr float PIO2_1 = 1.57079637050628662f;        // fl32(pi/2)
    constexpr float PIO2_2 = -4.37113882867379300e-08f;   // fl32(pi/2 - PIO2_1)
    constexpr float PIO2_3 = -1.71512451000588190e-15f;   // fl32(pi/2 - PIO2_1 - PIO2_2)
    const float k = rintf(x * TWO_OVER_PI);
    float r = fmaf(k, -PIO2_1, x);
    r = fmaf(k, -PIO2_2, r);
    r = fmaf(k, -PIO2_3, r);
    const float r2 = r * r;
    // 1/3!, 1/5!, ... and 1/2!, 1/4!, ... rounded to FP32
    float ps = fmaf(r2, 2.75573188446287530e-06f, -1.98412701138295233e-04f);
    ps = fmaf(r2, ps, 8.33333376795053482e-03f);
    ps = fmaf(r2, ps, -1.66666671633720398e-01f);
    ps = fmaf(r * r2, ps, r);                                             // sin r
    float pm = fmaf(r2, -2.75573199814971304e-07f, 2.48015876422869042e-05f);
    pm = fmaf(r2, pm, -1.38888892251998186e-03f);
    pm = fmaf(r2, pm, 4.16666679084300995e-02f);
    pm = fmaf(r2, pm, -0.5f);
    pm = pm * r2;                                                         // cos r - 1
    const float pc = 1.0f + pm;
    const int q = (int)k & 3;
    float sv = (q & 1) ? pc : ps, cv = (q & 1) ? ps : pc;
    if (q & 2) sv = -sv;
    if ((q + 1) & 2) cv = -cv;
    return {sv, cv, q == 0 ? pm : cv - 1.0f};
}

// FP64 sine and cosine of a link or plaquette angle, by the same scheme:
// pi/2 in three FP64 parts (FMA reduction), the Taylor series of sin r to
// r^17 and of cos r to r^16 on |r| <= pi/4 (truncation below 1e-19 and 3e-18 of
// the value, under 2^-53), Horner with FMAs, the quadrant k mod 4. Measured
// against 60-digit references: sin within 1.04 ulp, cos within 1.18 ulp (the
// device library's are of the same order; glibc's, which the reference uses,
// within 0.5), for |x| <= 13. Like any device sin/cos it is not glibc's bit for
// bit, which the gauged FP64 parity bar (1e-12 relative) allows for.
struct SinCos64 { double s, c; };
__device__ inline SinCos64 sincos64(double x) {
    constexpr double TWO_OVER_PI = 0.6366197723675814;
    constexpr double PIO2_1 = 1.5707963267948966;        // fl64(pi/2)
    constexpr double PIO2_2 = 6.123233995736766e-17;     // fl64(pi/2 - PIO2_1)
    constexpr double PIO2_3 = -1.4973849048591698e-33;   // fl64(pi/2 - PIO2_1 - PIO2_2)
    const double k = rint(x * TWO_OVER_PI);
    double r = fma(k, -PIO2_1, x);
    r = fma(k, -PIO2_2, r);
    r = fma(k, -PIO2_3, r);
    const double r2 = r * r;
    // (-1)^n / (2n+1)! and (-1)^n / (2n)!, n >= 1, rounded to FP64
    double ps = fma(r2, 2.8114572543455206e-15, -7.647163731819816e-13);
    ps = fma(r2, ps, 1.6059043836821613e-10);
    ps = fma(r2, ps, -2.505210838544172e-08);
    ps = fma(r2, ps, 2.7557319223985893e-06);
    ps = fma(r2, ps, -1.984126984126984e-04);
    ps = fma(r2, ps, 8.333333333333333e-03);
    ps = fma(r2, ps, -1.6666666666666666e-01);
    ps = fma(r * r2, ps, r);                                              // sin r
    double pm = fma(r2, 4.779477332387385e-14, -1.1470745597729725e-11);
    pm = fma(r2, pm, 2.08767569878681e-09);
    pm = fma(r2, pm, -2.755731922398589e-07);
    pm = fma(r2, pm, 2.48015873015873e-05);
    pm = fma(r2, pm, -1.388888888888889e-03);
    pm = fma(r2, pm, 4.1666666666666664e-02);
    pm = fma(r2, pm, -0.5);
    const double pc = fma(pm, r2, 1.0);                                   // cos r
    const int q = (int)k & 3;
    double sv = (q & 1) ? pc : ps, cv = (q & 1) ? ps : pc;
    if (q & 2) sv = -sv;
    if ((q + 1) & 2) cv = -cv;
    return {sv, cv};
}

// sin, cos (and cos - 1 for the FP32 difference forms; 0 in FP64, which does
// not use it) of a link angle
__device__ inline void link_sincos(double x, double &s, double &c, double &cm1) {
    const SinCos64 t = sincos64(x);
    s = t.s; c = t.c; cm1 = 0.0;
}
__device__ inline void link_sincos(float x, float &s, float &c, float &cm1) {
    const SinCos32 t = sincos32(x);
    s = t.s; c = t.c; cm1 = t.cm1;
}
__device__ inline double scp_sin(double x) { return sincos64(x).s; }
__device__ inline float scp_sin(float x) { return sincos32(x).s; }

constexpr int TK = 64, TJ = 4;

struct Geo {
    int N, i0, i1, j, k;
    bool valid;
};

// XCD-aware block numbering: hardware block b runs on XCD b % 8; give each XCD
// a contiguous range of logical tiles so j-adjacent tiles share an L2.
__device__ inline uint32_t logical_block(uint32_t b, uint32_t nb) {
    const uint32_t nx = 8;
    uint32_t per = (nb + nx - 1) / nx;
    uint32_t xcd = b % nx, pos = b / nx;
    uint32_t l = xcd * per + pos;
    // ragged tail: fall back to identity for the few blocks past the balanced part
    return (nb % nx == 0) ? l : b;
}

__device__ inline Geo geometry(int N, int DI) {
    Geo g;
    g.N = N;
    const uint32_t ntk = (N + TK - 1) / TK, ntj = (N + TJ - 1) / TJ;
    const uint32_t nb = gridDim.x;
    uint32_t lb = logical_block(blockIdx.x, nb);
    uint32_t tk = lb % ntk; lb /= ntk;
    uint32_t tj = lb % ntj; lb /= ntj;
    uint32_t seg = lb;
    g.k = tk * TK + threadIdx.x;
    g.j = tj * TJ + threadIdx.y;
    g.i0 = seg * DI;
    g.i1 = min(N, g.i0 + DI);
    g.valid = g.k < N && g.j < N;
    return g;
}

__device__ inline uint32_t sidx(int N, int i, int j, int k) {
    return ((uint32_t)i * (uint32_t)N + (uint32_t)j) * (uint32_t)N + (uint32_t)k;
}
__device__ inline int wp(int a, int N) { return a + 1 == N ? 0 : a + 1; }
__device__ inline int wm(int a, int N) { return a == 0 ? N - 1 : a - 1; }


// ---------------------------------------------------------------------------
// per-site physics (expression trees of scp_sim.c, kept operation for operation)
// ---------------------------------------------------------------------------

// sin of the plaquette angle th_a(y) + th_b(y+a) - th_a(y+b) - th_b(y)  (C:1100-1102)
template <typename R>
__device__ inline R plaq(R ta_y, R tb_ya, R ta_yb, R tb_y) {
    R ang = ta_y + tb_ya - ta_yb - tb_y;
    return scp_sin(ang);
}


// In-plane neighbour access for the per-site physics. A neighbour provider NB
// has face(f, d, sg): field f at the site one step along d (1 = j, 2 = k) in
// direction sg; and the four link diagonals the staples need:
//   th1_jm_kp: th_1(i, j-1, k+1)   th2_jp_km: th_2(i, j+1, k-1)
//   th1_ip_jm: th_1(i+1, j-1, k)   th2_ip_km: th_2(i+1, j, k-1)
// GlobalNB reads them from the device arrays; the LDS-staged kernels read them
// from the staged plane tiles.
template <typename R, class LD>
struct GlobalNB {
    LD X;
    int N, i, j, k;
    uint32_t jp, jm, kp, km;
    uint32_t rs, ps;   // row and plane strides of the source array
    __device__ uint32_t at(int a, int b, int c) const { return (uint32_t)a * ps + (uint32_t)b * rs + (uint32_t)c; }
    __device__ R face(int f, int d, int sg) const {
        return X(f, d == 1 ? (sg > 0 ? jp : jm) : (sg > 0 ? kp : km));
    }
    __device__ R th1_jm_kp() const { return X(F_TH + 1, at(i, wm(j, N), wp(k, N))); }
    __device__ R th2_jp_km() const { return X(F_TH + 2, at(i, wp(j, N), wm(k, N))); }
    __device__ R th1_ip_jm() const { return X(F_TH + 1, at(wp(i, N), wm(j, N), k)); }
    __device__ R th2_ip_km() const { return X(F_TH + 2, at(wp(i, N), j, wm(k, N))); }
};

// Complex-mode accelerations at one site (compute_forces_complex, C:931-994).
// c/ip/im: own column at planes i, i+1, i-1 (anything indexable by field);
// nb: in-plane neighbours.
template <typename R, class CV, class PV, class MV, class NB>
__device__ inline void forces_complex(const KParams &kp, const CV &c, const PV &ip,
                                      const MV &im, const NB &nb, R *acc) {
    const R h1 = R(kp.h1), h2 = R(kp.h2);
    R s2[3];
#pragma unroll
    for (int a = 0; a < 3; a++) s2[a] = c[F_U + a] * c[F_U + a] + c[F_V + a] * c[F_V + a];
    R s = s2[0] * s2[1] * s2[2];
    R den = R(1.0) + R(kp.KAPPA) * s;
    R Vp = R(0.5) * R(kp.MU) / (den * den);
    R rest[3] = {s2[1] * s2[2], s2[0] * s2[2], s2[0] * s2[1]};
    // neighbour value of field f in direction d (+1/-1)
    auto nbv = [&](int f, int d, int sg) -> R {
        if (d == 0) return sg > 0 ? ip[f] : im[f];
        return nb.face(f, d, sg);
    };
    auto lap = [&](int f) -> R {
        if constexpr (diff_form<R>()) {
            const R x = c[f];
            return (((nbv(f, 0, 1) - x) + (nbv(f, 0, -1) - x)) + ((nbv(f, 1, 1) - x) + (nbv(f, 1, -1) - x)) +
                    ((nbv(f, 2, 1) - x) + (nbv(f, 2, -1) - x))) * h2;
        }
        return (nbv(f, 0, 1) + nbv(f, 0, -1) + nbv(f, 1, 1) + nbv(f, 1, -1) + nbv(f, 2, 1) +
                nbv(f, 2, -1) - R(6.0) * c[f]) * h2;
    };
    // curl of the 3-vector at base f0, component a (C:409-415)
    auto curl = [&](int f0, int a) -> R {
        if (a == 0)
            return (nbv(f0 + 2, 1, 1) - nbv(f0 + 2, 1, -1) - nbv(f0 + 1, 2, 1) + nbv(f0 + 1, 2, -1)) * h1;
        if (a == 1)
            return (nbv(f0 + 0, 2, 1) - nbv(f0 + 0, 2, -1) - nbv(f0 + 2, 0, 1) + nbv(f0 + 2, 0, -1)) * h1;
        return (nbv(f0 + 1, 0, 1) - nbv(f0 + 1, 0, -1) - nbv(f0 + 0, 1, 1) + nbv(f0 + 0, 1, -1)) * h1;
    };
#pragma unroll
    for (int a = 0; a < 3; a++) {
        R u = c[F_U + a], v = c[F_V + a];
        acc[F_U + a] = lap(F_U + a) - R(kp.MASS2) * u - R(2.0) * Vp * u * rest[a] + R(kp.ETA) * curl(F_TU, a);
        acc[F_V + a] = lap(F_V + a) - R(kp.MASS2) * v - R(2.0) * Vp * v * rest[a] + R(kp.ETA) * curl(F_TV, a);
        acc[F_TU + a] = lap(F_TU + a) - R(kp.MT2) * c[F_TU + a] + R(kp.ETA) * curl(F_U, a);
        acc[F_TV + a] = lap(F_TV + a) - R(kp.MT2) * c[F_TV + a] + R(kp.ETA) * curl(F_V, a);
    }
}

// Real-mode accelerations (compute_forces, C:743-922). c/ip/im: own column at
// planes i, i+1, i-1; nb: in-plane neighbours; nbi(d, sg): the index of the
// neighbour along d (0 = i, 1 = j, 2 = k) in direction sg, nbi(3, 0) the
// site's own (called only where a pass-1 term is on). AUX(f, idx) loads the
// pass-1 fields M[0..2] (f=0..2) and Q[0..2] (f=3..5) when alpha_cs/beta_h != 0,
// and the local theta frequency (f=6) when sigma_freq != 0. With kp.XT the
// terms of the reference's GPU step follow (compute_forces_kernel,
// scp_sim.cu:1144-1248), added in its order onto the base accelerations.
template <typename R, class CV, class PV, class MV, class NB, class LA, class IX>
__device__ inline void forces_real(const KParams &kp, const CV &c, const PV &ip,
                                   const MV &im, const NB &nb, LA AUX, IX nbi, R *acc) {
    const R h1 = R(kp.h1), h2 = R(kp.h2);
    const R MU = R(kp.MU), KAPPA = R(kp.KAPPA), MASS2 = R(kp.MASS2), ETA = R(kp.ETA);
    const R KH = R(kp.KH), ACS = R(kp.ACS), BH = R(kp.BH), KG = R(kp.KG);
    const int MODE = kp.rmode;
    auto nbv = [&](int f, int d, int sg) -> R {
        if (d == 0) return sg > 0 ? ip[f] : im[f];
        return nb.face(f, d, sg);
    };
    auto curlF = [&](int f0, int a) -> R {
        if (a == 0)
            return (nbv(f0 + 2, 1, 1) - nbv(f0 + 2, 1, -1) - nbv(f0 + 1, 2, 1) + nbv(f0 + 1, 2, -1)) * h1;
        if (a == 1)
            return (nbv(f0 + 0, 2, 1) - nbv(f0 + 0, 2, -1) - nbv(f0 + 2, 0, 1) + nbv(f0 + 2, 0, -1)) * h1;
        return (nbv(f0 + 1, 0, 1) - nbv(f0 + 1, 0, -1) - nbv(f0 + 0, 1, 1) + nbv(f0 + 0, 1, -1)) * h1;
    };
    auto curlA = [&](int f0, int a) -> R {
        if (a == 0)
            return (AUX(f0 + 2, nbi(1, 1)) - AUX(f0 + 2, nbi(1, -1)) - AUX(f0 + 1, nbi(2, 1)) + AUX(f0 + 1, nbi(2, -1))) * h1;
        if (a == 1)
            return (AUX(f0 + 0, nbi(2, 1)) - AUX(f0 + 0, nbi(2, -1)) - AUX(f0 + 2, nbi(0, 1)) + AUX(f0 + 2, nbi(0, -1))) * h1;
        return (AUX(f0 + 1, nbi(0, 1)) - AUX(f0 + 1, nbi(0, -1)) - AUX(f0 + 0, nbi(1, 1)) + AUX(f0 + 0, nbi(1, -1))) * h1;
    };
    const int TU = 3;   // real mode: tu at fields 3..5
    R p0 = c[0], p1 = c[1], p2 = c[2];
    R sig = p0 * p0 + p1 * p1 + p2 * p2;
    R me2 = (MODE == 1) ? R(kp.IA) / (R(1.0) + R(kp.IB) * sig) : MASS2;
    R keff = (MODE == 3) ? KAPPA / (R(1.0) + KG * sig) : KAPPA;
    R P = p0 * p1 * p2, P2 = P * P;
    R den = R(1.0) + keff * P2;
    R dVdP = MU * P / (den * den);
    R t2c = 0;
    if (MODE == 3 && KG > 0) {
        R D = R(1.0) + KG * sig + KAPPA * P2;
        t2c = MU * KG * KAPPA * P2 * P2 / (D * D);
    }
    R cph[3] = {curlF(0, 0), curlF(0, 1), curlF(0, 2)};
    R hel = p0 * cph[0] + p1 * cph[1] + p2 * cph[2];
    R dP2[3] = {0, 0, 0};
    if (KH != 0) {
#pragma unroll
        for (int d = 0; d < 3; d++) {
            R Pp = nbv(0, d, 1) * nbv(1, d, 1) * nbv(2, d, 1);
            R Pm = nbv(0, d, -1) * nbv(1, d, -1) * nbv(2, d, -1);
            dP2[d] = (Pp * Pp - Pm * Pm) * h1;
        }
    }
    R cM[3] = {0, 0, 0}, cQ[3] = {0, 0, 0};
    if (ACS != 0) { cM[0] = curlA(0, 0); cM[1] = curlA(0, 1); cM[2] = curlA(0, 2); }
    if (BH != 0) { cQ[0] = curlA(3, 0); cQ[1] = curlA(3, 1); cQ[2] = curlA(3, 2); }
    R csq = cph[0] * cph[0] + cph[1] * cph[1] + cph[2] * cph[2];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        R lp;
        if constexpr (diff_form<R>()) {
            const R x = c[a];
            lp = (((nbv(a, 0, 1) - x) + (nbv(a, 0, -1) - x)) + ((nbv(a, 1, 1) - x) + (nbv(a, 1, -1) - x)) +
                  ((nbv(a, 2, 1) - x) + (nbv(a, 2, -1) - x))) * h2;
        } else {
            lp = (nbv(a, 0, 1) + nbv(a, 0, -1) + nbv(a, 1, 1) + nbv(a, 1, -1) + nbv(a, 2, 1) +
                  nbv(a, 2, -1) - R(6.0) * c[a]) * h2;
        }
        R dPda = (a == 0) ? p1 * p2 : (a == 1) ? p0 * p2 : p0 * p1;
        R ct = curlF(TU, a);
        R chiral = 0;
        if (KH != 0) {
            R t1 = R(2.0) * P * dPda * hel;
            R t2 = R(2.0) * P2 * cph[a];
            R t3;
            if (a == 0) t3 = -(dP2[2] * p1 - dP2[1] * p2);
            else if (a == 1) t3 = -(dP2[0] * p2 - dP2[2] * p0);
            else t3 = -(dP2[1] * p0 - dP2[0] * p1);
            chiral = KH * (t1 + t2 + t3);
        }
        R cos_phi = -ACS * cM[a];
        R hard_phi = -R(2.0) * cQ[a];
        acc[a] = lp - me2 * c[a] - dVdP * dPda - t2c * c[a] + ETA * ct + chiral + cos_phi + hard_phi;
    }
#pragma unroll
    for (int a = 0; a < 3; a++) {
        R lt;
        if constexpr (diff_form<R>()) {
            const R x = c[TU + a];
            lt = (((nbv(TU + a, 0, 1) - x) + (nbv(TU + a, 0, -1) - x)) +
                  ((nbv(TU + a, 1, 1) - x) + (nbv(TU + a, 1, -1) - x)) +
                  ((nbv(TU + a, 2, 1) - x) + (nbv(TU + a, 2, -1) - x))) * h2;
        } else {
            lt = (nbv(TU + a, 0, 1) + nbv(TU + a, 0, -1) + nbv(TU + a, 1, 1) + nbv(TU + a, 1, -1) +
                  nbv(TU + a, 2, 1) + nbv(TU + a, 2, -1) - R(6.0) * c[TU + a]) * h2;
        }
        R cp = curlF(0, a);
        R cos_th = 0;
        if (ACS != 0) cos_th = R(2.0) * ACS * AUX(a, nbi(3, 0));
        R hard_th = -BH * csq * c[TU + a];
        acc[TU + a] = lt - R(kp.MT2) * c[TU + a] + ETA * cp + cos_th + hard_th;
    }
    if (kp.XT) {
        const R th[3] = {c[TU], c[TU + 1], c[TU + 2]};
        const R ct[3] = {curlF(TU, 0), curlF(TU, 1), curlF(TU, 2)};
        const R th2 = th[0] * th[0] + th[1] * th[1] + th[2] * th[2];
        if (kp.TSAT > 0 && kp.GCONV > 0) {          // saturation -> conversion (CU:1144-1161)
            const R excess = th2 - R(kp.TSAT) * R(kp.TSAT);
            if (excess > 0) {
#pragma unroll
                for (int a = 0; a < 3; a++) {
                    acc[TU + a] -= R(kp.GCONV) * excess * th[a];
                    acc[a] += R(kp.GCONV) * excess * ct[a];
                }
            }
        }
        if (kp.SGRAD > 0) {                         // gradient self-interaction (CU:1163-1185)
            R g2 = 0;
#pragma unroll
            for (int a = 0; a < 3; a++) {
                const R dx_ = (nbv(TU + a, 0, 1) - nbv(TU + a, 0, -1)) * h1;
                const R dy_ = (nbv(TU + a, 1, 1) - nbv(TU + a, 1, -1)) * h1;
                const R dz_ = (nbv(TU + a, 2, 1) - nbv(TU + a, 2, -1)) * h1;
                g2 += dx_ * dx_ + dy_ * dy_ + dz_ * dz_;
            }
            const R hh = th[0] * ct[0] + th[1] * ct[1] + th[2] * ct[2];
            R se = R(kp.SGRAD) * (R(1.0) + R(kp.CHI) * hh / (th2 + R(0.01)));
            if (se < 0) se = 0;
#pragma unroll
            for (int a = 0; a < 3; a++) {
                acc[TU + a] -= se * g2 * th[a];
                acc[a] += se * g2 * ct[a];
            }
        }
        if (kp.SCUB > 0) {                          // cubic self-interaction (CU:1188-1199)
#pragma unroll
            for (int a = 0; a < 3; a++) {
                acc[TU + a] -= R(kp.SCUB) * th2 * th[a];
                acc[a] += R(kp.SCUB) * th2 * ct[a];
            }
        }
        if (kp.SFREQ > 0) {                         // frequency mismatch (CU:1201-1229)
            const R oc = AUX(6, nbi(3, 0));
            R dw = 0;
#pragma unroll
            for (int d = 0; d < 3; d++)
#pragma unroll
                for (int sg = 1; sg >= -1; sg -= 2) dw += fabs(oc - AUX(6, nbi(d, sg)));
            const R dwa = dw / R(6);
#pragma unroll
            for (int a = 0; a < 3; a++) {
                acc[TU + a] -= R(kp.SFREQ) * dwa * th[a];
                acc[a] += R(kp.SFREQ) * dwa * ct[a];
            }
        }
        if (kp.SCROSS > 0) {                        // cross potential (CU:1232-1240)
            const R ph2 = c[0] * c[0] + c[1] * c[1] + c[2] * c[2];
#pragma unroll
            for (int a = 0; a < 3; a++) {
                acc[a] -= R(kp.SCROSS) * th2 * c[a];
                acc[TU + a] -= R(kp.SCROSS) * ph2 * th[a];
            }
        }
        if (kp.LSELF > 0) {                         // theta self potential (CU:1243-1248)
#pragma unroll
            for (int a = 0; a < 3; a++) acc[TU + a] -= R(kp.LSELF) * th2 * th[a];
        }
    }
}

// Carried per-column gauge quantities from plane i-1.
template <typename R>
struct GaugeCarry {
    R c0, s0, m0;  // cos, sin, cos - 1 of th_0(x - e0)
    R p0, p2;           // plaquette sines (0,1) and (2,0) at x - e0
};

// Gauged accelerations (compute_forces_complex_gauge, C:1053-1263), one
// gauge, no Higgs. Writes acc[0..11] (matter, device order) and acc[12..14]
// (E_acc), and the carry for plane i+1.
template <typename R, class CV, class PV, class MV, class NB>
__device__ inline void forces_gauge(const KParams &kp, const CV &c, const PV &ip,
                                    const MV &im, const NB &nb, GaugeCarry<R> &cy, R *acc) {
    const R h1 = R(kp.h1), h2 = R(kp.h2);
    auto nbv = [&](int f, int d, int sg) -> R {
        if (d == 0) return sg > 0 ? ip[f] : im[f];
        return nb.face(f, d, sg);
    };
    // links and their cos/sin at x and x - e_d (C:1071-1079, 1144-1145)
    // (mP, mM: cos - 1, read by the FP32 difference forms only)
    R cP[3], sP[3], cM[3], sM[3], mP[3], mM[3];
#pragma unroll
    for (int d = 0; d < 3; d++) link_sincos(c[F_TH + d], sP[d], cP[d], mP[d]);
    cM[0] = cy.c0; sM[0] = cy.s0; mM[0] = cy.m0;
    link_sincos(nbv(F_TH + 1, 1, -1), sM[1], cM[1], mM[1]);
    link_sincos(nbv(F_TH + 2, 2, -1), sM[2], cM[2], mM[2]);

    // plaquettes (C:1088-1109) at x, x-e1, x-e2; x-e0 carried
    R th0 = c[F_TH + 0], th1 = c[F_TH + 1], th2 = c[F_TH + 2];
    R px0 = plaq(th0, ip[F_TH + 1], nbv(F_TH + 0, 1, 1), th1);       // (0,1) at x
    R px1 = plaq(th1, nbv(F_TH + 2, 1, 1), nbv(F_TH + 1, 2, 1), th2); // (1,2) at x
    R px2 = plaq(th2, nbv(F_TH + 0, 2, 1), ip[F_TH + 2], th0);       // (2,0) at x
    // y = x - e1
    R t1_ym = nbv(F_TH + 1, 1, -1), t0_ym = nbv(F_TH + 0, 1, -1), t2_ym = nbv(F_TH + 2, 1, -1);
    R p0_ym = plaq(t0_ym, nb.th1_ip_jm(), th0, t1_ym);
    R p1_ym = plaq(t1_ym, th2, nb.th1_jm_kp(), t2_ym);
    // y = x - e2
    R t1_zm = nbv(F_TH + 1, 2, -1), t2_zm = nbv(F_TH + 2, 2, -1), t0_zm = nbv(F_TH + 0, 2, -1);
    R p1_zm = plaq(t1_zm, nb.th2_jp_km(), th1, t2_zm);
    R p2_zm = plaq(t2_zm, th0, nb.th2_ip_km(), t0_zm);
    R st[3];
    st[0] = (px0 - p0_ym) - (px2 - p2_zm);
    st[1] = -(px0 - cy.p0) + (px1 - p1_zm);
    st[2] = (px2 - cy.p2) - (px1 - p1_ym);

    // potential (C:1160-1166)
    R fu[6], fv[6];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        fu[a] = c[F_U + a]; fv[a] = c[F_V + a];
        fu[3 + a] = c[F_TU + a]; fv[3 + a] = c[F_TV + a];
    }
    R s2[3];
#pragma unroll
    for (int a = 0; a < 3; a++) s2[a] = fu[a] * fu[a] + fv[a] * fv[a];
    R s = s2[0] * s2[1] * s2[2];
    R den = R(1.0) + R(kp.KAPPA) * s;
    R Vp = R(0.5) * R(kp.MU) / (den * den);
    R rest[3] = {s2[1] * s2[2], s2[0] * s2[2], s2[0] * s2[1]};

    // transported neighbours, per direction (C:1134-1157), consumed as produced
    R lU[6], lV[6];          // Laplacian partial sums, per field f (Phi 0..2, Theta 0..2)
    R DU[3][6], DV[3][6];    // covariant central differences (only d != component used)
    const int C1[3] = {1, 2, 0}, C2[3] = {2, 0, 1};
#pragma unroll
    for (int d = 0; d < 3; d++) {
        R TRp[6], TIp[6], tup[3], tvp[3];
        R Jg = R(0.0);
        // FP32: transported differences U x_n - x (see diff_form)
        const R cmP = diff_form<R>() ? mP[d] : R(0);
        const R cmM = diff_form<R>() ? mM[d] : R(0);
#pragma unroll
        for (int a = 0; a < 3; a++) {
#pragma unroll
            for (int h = 0; h < 2; h++) {       // h=0: Phi_a (u,v); h=1: Theta_a (tu,tv)
                const int fuix = h ? F_TU + a : F_U + a, fvix = h ? F_TV + a : F_V + a;
                R upn = nbv(fuix, d, 1), vpn = nbv(fvix, d, 1);
                R umn = nbv(fuix, d, -1), vmn = nbv(fvix, d, -1);
                const int f = h * 3 + a;
                if constexpr (diff_form<R>()) {
                    const R u = fu[f], v = fv[f];
                    const R epr = (upn - u) + (cmP * upn - sP[d] * vpn);   // trp - u
                    const R epi = (vpn - v) + (cmP * vpn + sP[d] * upn);   // tip - v
                    const R emr = (umn - u) + (cmM * umn + sM[d] * vmn);   // trm - u
                    const R emi = (vmn - v) + (cmM * vmn - sM[d] * umn);   // tim - v
                    if (d == 0) { lU[f] = epr + emr; lV[f] = epi + emi; }
                    else { lU[f] = lU[f] + (epr + emr); lV[f] = lV[f] + (epi + emi); }
                    DU[d][f] = (epr - emr) * h1;
                    DV[d][f] = (epi - emi) * h1;
                    TRp[f] = u + epr; TIp[f] = v + epi;
                    Jg += u * epi - v * epr;                               // = u tip - v trp
                } else {
                    R trp = cP[d] * upn - sP[d] * vpn, tip = cP[d] * vpn + sP[d] * upn;
                    R trm = cM[d] * umn + sM[d] * vmn, tim = cM[d] * vmn - sM[d] * umn;
                    if (d == 0) { lU[f] = trp + trm; lV[f] = tip + tim; }
                    else { lU[f] = lU[f] + trp; lU[f] = lU[f] + trm; lV[f] = lV[f] + tip; lV[f] = lV[f] + tim; }
                    DU[d][f] = (trp - trm) * h1;
                    DV[d][f] = (tip - tim) * h1;
                    TRp[f] = trp; TIp[f] = tip;
                }
                if (h) { tup[a] = upn; tvp[a] = vpn; }
            }
        }
        // lattice current J_d and E-kick (C:1232-1250)
        if constexpr (!diff_form<R>()) {
#pragma unroll
            for (int a = 0; a < 3; a++) {
                Jg += fu[a] * TIp[a] - fv[a] * TRp[a];
                Jg += fu[3 + a] * TIp[3 + a] - fv[3 + a] * TRp[3 + a];
            }
        }
        const int d1 = C1[d], d2 = C2[d];
        R WpR1 = cP[d] * fu[d2] + sP[d] * fv[d2], WpI1 = cP[d] * fv[d2] - sP[d] * fu[d2];
        R WpR2 = cP[d] * fu[d1] + sP[d] * fv[d1], WpI2 = cP[d] * fv[d1] - sP[d] * fu[d1];
        R T12 = (fu[3 + d1] * TIp[d2] - fv[3 + d1] * TRp[d2]) + (tup[d1] * WpI1 - tvp[d1] * WpR1);
        R T21 = (fu[3 + d2] * TIp[d1] - fv[3 + d2] * TRp[d1]) + (tup[d2] * WpI2 - tvp[d2] * WpR2);
        R Jlat = R(kp.inva) * Jg - R(0.5) * R(kp.ETA) * (T12 - T21);
        acc[F_TH + d] = R(kp.STP) * st[d] + R(kp.GG) * Jlat;
    }
    // matter accelerations (C:1179-1204)
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const int d1 = C1[a], d2 = C2[a];
        R reDxT = DU[d1][3 + C2[a]] - DU[d2][3 + C1[a]];
        R imDxT = DV[d1][3 + C2[a]] - DV[d2][3 + C1[a]];
        R reDxP = DU[d1][C2[a]] - DU[d2][C1[a]];
        R imDxP = DV[d1][C2[a]] - DV[d2][C1[a]];
        R lUP, lVP, lUT, lVT;
        if constexpr (diff_form<R>()) {     // lU, lV already hold sum_n (U x_n - x)
            lUP = lU[a] * h2; lVP = lV[a] * h2; lUT = lU[3 + a] * h2; lVT = lV[3 + a] * h2;
        } else {
            lUP = (lU[a] - R(6.0) * fu[a]) * h2;
            lVP = (lV[a] - R(6.0) * fv[a]) * h2;
            lUT = (lU[3 + a] - R(6.0) * fu[3 + a]) * h2;
            lVT = (lV[3 + a] - R(6.0) * fv[3 + a]) * h2;
        }
        acc[F_U + a] = lUP - R(kp.MASS2) * fu[a] - R(2.0) * Vp * fu[a] * rest[a] + R(kp.ETA) * reDxT;
        acc[F_V + a] = lVP - R(kp.MASS2) * fv[a] - R(2.0) * Vp * fv[a] * rest[a] + R(kp.ETA) * imDxT;
        acc[F_TU + a] = lUT - R(kp.MT2) * fu[3 + a] + R(kp.ETA) * reDxP;
        acc[F_TV + a] = lVT - R(kp.MT2) * fv[3 + a] + R(kp.ETA) * imDxP;
    }
    cy.c0 = cP[0]; cy.s0 = sP[0]; cy.m0 = mP[0]; cy.p0 = px0; cy.p2 = px2;
}

// carry for the segment's first plane, from the links at (i-1, j, k) (mt0..2),
// th_1, th_2 at (i, j, k) and th_0 at (i-1, j+1, k), (i-1, j, k+1)
template <typename R>
__device__ inline void gauge_prologue(R mt0, R mt1, R mt2, R th1, R th2, R t0_mjp, R t0_mkp,
                                      GaugeCarry<R> &cy) {
    link_sincos(mt0, cy.s0, cy.c0, cy.m0);
    cy.p0 = plaq(mt0, th1, t0_mjp, mt1);
    cy.p2 = plaq(mt2, t0_mkp, th2, mt0);
}

// sponge factor of scp_sim.c:1447-1453 (1.0 and false outside the sponge)
__device__ inline bool sponge(const KParams &kp, int i, int j, int k, double &d) {
    double x = -kp.L + i * kp.dx, y = -kp.L + j * kp.dx, z = -kp.L + k * kp.dx;
    double r = sqrt(x * x + y * y + z * z);
    if (r > kp.Rd) {
        double s = (r - kp.Rd) / kp.DW;
        if (s > 1) s = 1;
        d = 1.0 - kp.DR * s * s;
        return true;
    }
    return false;
}

}  // namespace scp
