// scp_abi.hip — the C-ABI of include/scp_amd.h over the HIP kernels.
//
// Device state of a grid (one allocation per group, fields 64-element aligned):
//   P[2]  positions, ping-pong (the stencil source and the drift target)
//   W     half-kicked velocities / E, updated in place
//   xo,vo positions and velocities of the last completed step (FP64), written
//         by the last launch of every scp_grid_step call and by upload
// See scp_step.hip for the step schedule.
#include "scp_grid.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

using namespace scp;

namespace scp_internal {


// K_STEP runs the LDS-staged march unless the caller asked for the
// direct-neighbour kernel; bc 1 (pinned slabs) runs the direct kernel. The
// march addresses a field with 32-bit byte offsets, so one field must stay
// below 4 GiB (N <= ~800 in FP64, N <= 1023 in FP32C); larger grids run the
// direct-neighbour kernel.
bool use_march(const scp_grid *g, int kind) {
    const uint64_t field_bytes = (uint64_t)g->kp.fsh * (g->f32 ? 4u : 8u);
    return kind == K_STEP && g->p.step_kernel == 0 && g->p.bc_type != 1 && field_bytes < (1ull << 32);
}

// the march instance of a grid: its step mode, or 3 for real-mode grids that
// carry the pass-1 fields (the instance that reads them)
static int march_mode(const scp_grid *g) { return g->step_mode == 0 && g->aux ? 3 : g->step_mode; }

static thread_local char g_err[512];

int set_err(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

int kerr(int rc, const char *what) {
    if (rc != 0) return set_err(-2, "%s launch failed: %s", what, hipGetErrorString((hipError_t)rc));
    return 0;
}

}  // namespace scp_internal

using namespace scp_internal;

extern "C" int scp_abi_version(void) { return SCP_ABI_VERSION; }

extern "C" int scp_last_error(char *buf, int n) {
    int len = (int)strlen(g_err);
    if (buf && n > 0) {
        strncpy(buf, g_err, (size_t)n - 1);
        buf[n - 1] = 0;
    }
    return len;
}

static int ncols_of(int mode) { return mode == 2 ? 30 : (mode == 1 ? 24 : 12); }

// buffers an entry point needs: the FP64 emit buffers / the step state (a
// pooled grid of the request loop holds the first only while it must emit, the
// second only until its last step)
static int need_emit(const scp_grid *g) {
    return (g->xo && g->vo) ? 0 : set_err(-1, "grid holds no FP64 emit buffers (released by its owner)");
}
static int need_state(const scp_grid *g) {
    return (g->P[0] || g->P32[0]) ? 0 : set_err(-1, "grid holds no step state (released after its last step)");
}


// device field f (position, or velocity when vel) -> SFA snapshot column
static int col_of(int mode, int f, bool vel) {
    if (mode == 0) return vel ? 6 + f : f;        // u0..2 -> 0..2, tu0..2 -> 3..5
    static const int POS[15] = {0, 1, 2, 12, 13, 14, 3, 4, 5, 15, 16, 17, 24, 25, 26};
    static const int VEL[15] = {6, 7, 8, 18, 19, 20, 9, 10, 11, 21, 22, 23, 27, 28, 29};
    return vel ? VEL[f] : POS[f];
}

// device field f -> reference acceleration column (phi_acc, theta_acc, phi_im_acc, theta_im_acc, E_acc)
static int acc_col_of(int mode, int f) {
    if (mode == 0) return f;
    static const int A[15] = {0, 1, 2, 6, 7, 8, 3, 4, 5, 9, 10, 11, 12, 13, 14};
    return A[f];
}

static uint32_t field_stride(int64_t n3) { return (uint32_t)((n3 + 63) / 64 * 64); }
// row pitch of the step state: 32-element rows (128 B in FP32) for the
// complex and gauged steps; the real step (direct kernel only) keeps N
static uint32_t step_pitch(int N, int mode) { return mode == 0 ? (uint32_t)N : (uint32_t)((N + 31) / 32 * 32); }
static uint32_t step_field_stride(int N, int mode) {
    return field_stride((int64_t)N * N * step_pitch(N, mode));
}

static int check_params(const scp_params *p) {
    if (!p) return set_err(-1, "null params");
    if (p->N < 4 || p->N > 1290) return set_err(-1, "N=%d out of range [4, 1290]", p->N);
    if (p->mode < 0 || p->mode > 2) return set_err(-1, "mode=%d invalid", p->mode);
    if (p->precision != SCP_PREC_F64 && p->precision != SCP_PREC_F32C)
        return set_err(-1, "precision=%d invalid", p->precision);
    if (p->precision == SCP_PREC_F32C && (p->lo_words < 0 || p->lo_words > 4))
        return set_err(-1, "lo_words=%d invalid", p->lo_words);
    if (p->step_kernel < 0 || p->step_kernel > 1) return set_err(-1, "step_kernel=%d invalid", p->step_kernel);
    if (p->bc_type < 0 || p->bc_type > 2) return set_err(-1, "bc_type=%d invalid", p->bc_type);
    if (p->bc_type == 1) {   // gradient pinned (apply_gradient_bc, scp_sim.c:1480-1555)
        if (p->mode == 2)     // as scp_config.h:492-494
            return set_err(-1, "bc_type=1 is incompatible with the gauged mode (pinning A is gauge-dependent)");
        if (p->precision != SCP_PREC_F64) return set_err(-3, "bc_type=1 runs on the FP64 path only");
        if (p->gradient_margin < 0) return set_err(-1, "gradient_margin=%d must be >= 0", p->gradient_margin);
    } else if (p->gradient_margin != 0) {
        return set_err(-1, "gradient_margin applies to bc_type=1 only (must be 0)");
    }
    if (!(p->dx > 0) || !(p->dt > 0)) return set_err(-1, "dx/dt must be positive");
    if (p->mode != 0 && (p->theta_sat != 0 || p->gamma_conv != 0 || p->sigma_grad != 0 || p->chi_chiral != 0 ||
                         p->sigma_cubic != 0 || p->sigma_freq != 0 || p->sigma_cross != 0 || p->lambda_self != 0))
        return set_err(-1, "theta_sat/gamma_conv/sigma_*/chi_chiral/lambda_self are real-mode terms "
                           "(scp_config.h:440-463 refuses them with complex_phi)");
    if (p->mode == 2 && p->g < 0) return set_err(-1, "g must be >= 0");
    return 0;
}

extern "C" int64_t scp_grid_bytes(const scp_params *p) {
    if (check_params(p)) return -1;
    int64_t n3 = (int64_t)p->N * p->N * p->N;
    int64_t fs = field_stride(n3);
    int NP = np_of_mode(p->mode);
    const int64_t fh = step_field_stride(p->N, p->mode);
    int64_t b = 2LL * NP * fs * 8;                       // xo, vo
    if (p->precision == SCP_PREC_F64) b += 3LL * NP * fh * 8;           // P[2], W
    else b += fh * (12LL * NP + lo_site_bytes(lo_kind_w(p->lo_words), NP) + 2LL * lo_site_bytes(lo_kind_p(p->lo_words), NP));
    if (p->mode == 0 && (p->alpha_cs != 0 || p->beta_h != 0 || p->sigma_freq > 0)) b += 7LL * fs * 8;
    if (p->bc_type == 1) b += 2LL * NP * fs * 8 + 6LL * NP * p->N * p->N * 8;   // pins, side rows
    b += 1024LL * 32 * 8 + 4096;
    return b;
}

static KParams make_kparams(const scp_params &p, int step_mode) {
    KParams k;
    memset(&k, 0, sizeof k);
    k.N = p.N;
    k.rmode = p.rmode;
    int64_t n3 = (int64_t)p.N * p.N * p.N;
    k.n3 = (uint32_t)n3;
    k.fs = field_stride(n3);
    k.pitch = step_pitch(p.N, p.mode);
    k.plane = (uint32_t)p.N * k.pitch;
    k.fsh = step_field_stride(p.N, p.mode);
    const double dx = p.dx, dt = p.dt;
    k.hdt = 0.5 * dt;                        // scp_sim.c:1584
    k.dt = dt;
    k.gad = -p.g * dx * dt;                  // scp_sim.c:1642
    k.h1 = 1.0 / (2.0 * dx);                 // scp_sim.c:935
    k.h2 = 1.0 / (dx * dx);                  // scp_sim.c:934
    k.inva = 1.0 / dx;                       // scp_sim.c:1059
    k.GG = p.g;
    k.STP = step_mode == 2 ? 1.0 / (p.g * dx * dx * dx) : 0.0;   // scp_sim.c:1063
    k.MU = p.mu; k.KAPPA = p.kappa; k.MASS2 = p.m2; k.MT2 = p.mtheta2; k.ETA = p.eta;
    k.KH = p.kappa_h; k.ACS = p.alpha_cs; k.BH = p.beta_h;
    k.IA = p.inv_alpha; k.IB = p.inv_beta; k.KG = p.kappa_gamma;
    k.TSAT = p.theta_sat; k.GCONV = p.gamma_conv; k.SGRAD = p.sigma_grad; k.CHI = p.chi_chiral;
    k.SCUB = p.sigma_cubic; k.SFREQ = p.sigma_freq; k.SCROSS = p.sigma_cross; k.LSELF = p.lambda_self;
    k.XT = (p.mode == 0 && ((p.theta_sat > 0 && p.gamma_conv > 0) || p.sigma_grad > 0 || p.sigma_cubic > 0 ||
                            p.sigma_freq > 0 || p.sigma_cross > 0 || p.lambda_self > 0)) ? 1 : 0;
    k.L = p.L; k.dx = dx;
    k.Rd = p.L - p.damp_width;               // scp_sim.c:1444
    k.DW = p.damp_width; k.DR = p.damp_rate;
    k.damp = (p.bc_type == 0 && p.damp_width > 0 && p.damp_rate > 0) ? 1 : 0;
    k.bc1 = p.bc_type == 1 ? 1 : 0;
    k.gm = p.bc_type == 1 ? p.gradient_margin : 0;
    return k;
}

hipError_t scp_internal::gmalloc(scp_grid *g, void **p, size_t bytes) {
    *p = nullptr;
    return g->mem ? g->mem->alloc(g->mem->ctx, p, bytes) : hipMalloc(p, bytes);
}

void scp_internal::gfree(scp_grid *g, void *p) {
    if (!p) return;
    if (g->mem) g->mem->release(g->mem->ctx, p);
    else hipFree(p);
}

template <typename T>
static hipError_t galloc(scp_grid *g, T **p, size_t bytes) {
    return gmalloc(g, reinterpret_cast<void **>(p), bytes);
}

template <typename T>
static void grelease(scp_grid *g, T *&p) {
    gfree(g, p);
    p = nullptr;
}

static void release_step_state(scp_grid *g) {
    grelease(g, g->P[0]); grelease(g, g->P[1]); grelease(g, g->W);
    grelease(g, g->P32[0]); grelease(g, g->P32[1]); grelease(g, g->W32);
    grelease(g, g->Plo[0]); grelease(g, g->Plo[1]); grelease(g, g->Wlo);
}

static void free_grid(scp_grid *g) {
    if (!g) return;
    release_step_state(g);
    grelease(g, g->xo); grelease(g, g->vo);
    grelease(g, g->aux); grelease(g, g->acc); grelease(g, g->partials); grelease(g, g->dout);
    grelease(g, g->argmin); grelease(g, g->rho); grelease(g, g->fpart);
    grelease(g, g->pinx); grelease(g, g->pinv); grelease(g, g->side);
    if (g->ev0) hipEventDestroy(g->ev0);
    if (g->ev1) hipEventDestroy(g->ev1);
    if (g->stream) hipStreamDestroy(g->stream);
    delete g;
}

int64_t scp_internal::grid_emit_bytes(const scp_params *p) {
    const int64_t n3 = (int64_t)p->N * p->N * p->N;
    return 2LL * np_of_mode(p->mode) * field_stride(n3) * 8;
}

int scp_internal::grid_emit_buffers(scp_grid *g, bool on) {
    if (!on) {
        HIPCHK(hipStreamSynchronize(g->stream));
        grelease(g, g->xo);
        grelease(g, g->vo);
        return 0;
    }
    if (g->xo && g->vo) return 0;
    const size_t grp = (size_t)g->kp.fs * sizeof(double) * g->NP;
    hipError_t e = g->xo ? hipSuccess : galloc(g, &g->xo, grp);
    if (e == hipSuccess && !g->vo) e = galloc(g, &g->vo, grp);
    if (e != hipSuccess) {
        grelease(g, g->xo);
        grelease(g, g->vo);
        return set_err(-2, "emit buffer allocation failed: %s", hipGetErrorString(e));
    }
    return 0;
}

int scp_internal::grid_release_step_state(scp_grid *g) {
    HIPCHK(hipStreamSynchronize(g->stream));
    release_step_state(g);
    return 0;
}

int scp_internal::grid_create_with(const scp_params *p, const GridMem *mem, bool emit, scp_grid **out) {
    if (!out) return set_err(-1, "null out");
    *out = nullptr;
    int rc = check_params(p);
    if (rc) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return set_err(-4, "no HIP device");
    scp_grid *g = new scp_grid;
    memset(g, 0, sizeof *g);
    g->mem = mem;
    g->p = *p;
    g->mode = p->mode;
    g->step_mode = (p->mode == 2 && p->g == 0.0) ? 1 : p->mode;
    g->NP = np_of_mode(p->mode);
    g->kp = make_kparams(*p, g->step_mode);
    hipGetDevice(&g->device);
    const size_t fsb = (size_t)g->kp.fs * sizeof(double);
    const size_t grp = fsb * g->NP;                                      // xo, vo
    const size_t grph = (size_t)g->kp.fsh * sizeof(double) * g->NP;      // P, W (padded rows)
    hipError_t e = hipSuccess;
    e = e ? e : hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking);
    g->f32 = p->precision == SCP_PREC_F32C;
    g->lo_kind = p->lo_words;
    if (!g->f32) {
        e = e ? e : galloc(g, &g->P[0], grph);
        e = e ? e : galloc(g, &g->P[1], grph);
        e = e ? e : galloc(g, &g->W, grph);
    } else {
        const size_t g32 = (size_t)g->kp.fsh * 4 * g->NP;
        const size_t glw = (size_t)g->kp.fsh * lo_site_bytes(lo_kind_w(g->lo_kind), g->NP);
        const size_t glp = (size_t)g->kp.fsh * lo_site_bytes(lo_kind_p(g->lo_kind), g->NP);
        e = e ? e : galloc(g, &g->P32[0], g32);
        e = e ? e : galloc(g, &g->P32[1], g32);
        e = e ? e : galloc(g, &g->W32, g32);
        if (glp) {
            e = e ? e : galloc(g, &g->Plo[0], glp);
            e = e ? e : galloc(g, &g->Plo[1], glp);
        }
        if (glw) e = e ? e : galloc(g, &g->Wlo, glw);
    }
    if (emit) {
        e = e ? e : galloc(g, &g->xo, grp);
        e = e ? e : galloc(g, &g->vo, grp);
    }
    if (p->bc_type == 1) {
        e = e ? e : galloc(g, &g->pinx, grp);
        e = e ? e : galloc(g, &g->pinv, grp);
        e = e ? e : galloc(g, &g->side, (size_t)6 * g->NP * p->N * p->N * sizeof(double));
    }
    if (p->mode == 0 && (p->alpha_cs != 0 || p->beta_h != 0 || p->sigma_freq > 0)) {
        e = e ? e : galloc(g, &g->aux, 7 * fsb);
        e = e ? e : hipMemsetAsync(g->aux, 0, 7 * fsb, g->stream);
    }
    g->nblocks_diag = (int)std::min<int64_t>(1024, ((int64_t)g->kp.n3 + DIAG_BLOCK - 1) / DIAG_BLOCK);
    e = e ? e : galloc(g, &g->partials, (size_t)g->nblocks_diag * 32 * sizeof(double));
    e = e ? e : galloc(g, &g->dout, 64 * sizeof(double));
    e = e ? e : galloc(g, &g->argmin, sizeof(unsigned long long));
    e = e ? e : hipEventCreate(&g->ev0);
    e = e ? e : hipEventCreate(&g->ev1);
    if (e != hipSuccess) {
        set_err(-2, "allocation failed: %s", hipGetErrorString(e));
        free_grid(g);
        return -2;
    }
    if (use_march(g, K_STEP)) {
        if (g->f32)
            snprintf(g->kname, sizeof g->kname, "k_march<scp::S32<%d>, 1, %d, false>", g->lo_kind, march_mode(g));
        else
            snprintf(g->kname, sizeof g->kname, "k_march<scp::S64, 1, %d, false>", march_mode(g));
    } else if (g->f32)
        snprintf(g->kname, sizeof g->kname, "k_step32<%d, 1, false, %d>", g->step_mode, g->lo_kind);
    else
        snprintf(g->kname, sizeof g->kname, "k_step64<%d, 1, false>", g->step_mode);
    *out = g;
    return 0;
}

extern "C" int scp_grid_create(const scp_params *p, scp_grid **out) { return grid_create_with(p, nullptr, true, out); }

extern "C" int scp_grid_destroy(scp_grid *g) {
    if (!g) return 0;
    if (g->batch) scp_batch_remove(g->batch, g);   // a destroyed grid leaves its batch
    hipStreamSynchronize(g->stream);
    free_grid(g);
    return 0;
}

Bufs<double> scp_internal::bufs_for(scp_grid *g, int kind) {
    Bufs<double> b;
    memset(&b, 0, sizeof b);
    b.xo = g->xo;
    b.vo = g->vo;
    b.w = g->W;
    b.aux = g->aux;
    b.acc = g->acc;
    b.pinx = g->pinx;
    b.pinv = g->pinv;
    b.side = g->side;
    if (kind == K_INIT) { b.pin = g->xo; b.pout = g->P[0]; }
    else if (kind == K_STEP) { b.pin = g->P[g->cur]; b.pout = g->P[g->cur ^ 1]; }
    else { b.pin = g->xo; }
    return b;
}

// real-mode pass 1 from FP64 positions and the velocities the force sees
static int run_aux(scp_grid *g, const double *src, const double *vel) {
    if (!g->aux) return 0;
    return kerr(launch_real_aux(g->kp, src, vel, g->aux, g->stream), "real aux");
}

Bufs32 scp_internal::bufs32_for(scp_grid *g, int kind) {
    Bufs32 b;
    memset(&b, 0, sizeof b);
    b.xo = g->xo; b.vo = g->vo; b.w = g->W32; b.wlo = g->Wlo; b.aux = g->aux; b.acc = g->acc;
    if (kind == K_INIT) { b.pout = g->P32[0]; b.poutlo = g->Plo[0]; }
    else if (kind == K_STEP) {
        b.pin = g->P32[g->cur]; b.pinlo = g->Plo[g->cur];
        b.pout = g->P32[g->cur ^ 1]; b.poutlo = g->Plo[g->cur ^ 1];
    }
    return b;
}

static int launch_kind(scp_grid *g, int kind, bool emit) {
    if (use_march(g, kind)) {
        if (g->f32)
            return kerr(launch_march_f32c(march_mode(g), emit, g->lo_kind, g->kp, bufs32_for(g, kind), 0, g->stream),
                        "march32");
        return kerr(launch_march_f64(march_mode(g), emit, g->kp, bufs_for(g, kind), 0, g->stream), "march64");
    }
    if (g->f32)
        return kerr(launch_step_f32c(g->step_mode, kind, emit, g->lo_kind, g->kp, bufs32_for(g, kind), g->stream), "step32");
    return kerr(launch_step_f64(g->step_mode, kind, emit, g->kp, bufs_for(g, kind), g->stream), "step64");
}

// initial force + first half kick + first drift, from xo/vo; pin: bc 1 saves the
// state as the pinned state (an initial state; not after a rescale between steps)
int scp_internal::run_init(scp_grid *g, bool pin) {
    int rc = run_aux(g, g->xo, g->vo);
    if (rc) return rc;
    rc = launch_kind(g, K_INIT, false);
    if (rc) return rc;
    g->cur = 0;
    if (g->pinx && pin) {   // bc 1: the state at init is the pinned state (grid_save_pinned, C:381-400, 2845-2849)
        const size_t nb = (size_t)g->NP * g->kp.fs * sizeof(double);
        HIPCHK(hipMemcpyAsync(g->pinx, g->xo, nb, hipMemcpyDeviceToDevice, g->stream));
        HIPCHK(hipMemcpyAsync(g->pinv, g->vo, nb, hipMemcpyDeviceToDevice, g->stream));
    }
    return 0;
}

extern "C" int scp_grid_upload(scp_grid *g, const double *const *cols, int ncols) {
    if (!g || !cols) return set_err(-1, "null argument");
    if (need_emit(g)) return -1;
    if (need_state(g)) return -1;
    if (ncols != ncols_of(g->mode)) return set_err(-1, "ncols=%d, mode %d needs %d", ncols, g->mode, ncols_of(g->mode));
    const size_t nb = (size_t)g->kp.n3 * sizeof(double);
    for (int f = 0; f < g->NP; f++) {
        const double *xp = cols[col_of(g->mode, f, false)], *vp = cols[col_of(g->mode, f, true)];
        if (!xp || !vp) return set_err(-1, "null column");
        HIPCHK(hipMemcpyAsync(g->xo + (size_t)f * g->kp.fs, xp, nb, hipMemcpyHostToDevice, g->stream));
        HIPCHK(hipMemcpyAsync(g->vo + (size_t)f * g->kp.fs, vp, nb, hipMemcpyHostToDevice, g->stream));
    }
    int rc = run_init(g);
    if (rc) return rc;
    g->steps = 0;
    HIPCHK(hipStreamSynchronize(g->stream));
    return 0;
}

extern "C" int scp_grid_seed_balls(scp_grid *g, const scp_qball *balls, int nballs, double theta_sigma,
                                   double link_sigma, uint64_t seed) {
    if (!g) return set_err(-1, "null grid");
    if (g->mode == 0) return set_err(-1, "seeding Q-balls needs the complex or gauged mode");
    if (nballs < 0 || (nballs > 0 && !balls)) return set_err(-1, "bad ball list");
    for (int b = 0; b < nballs; b++)
        if (!(balls[b].sigma > 0) || (balls[b].winding != 1 && balls[b].winding != -1))
            return set_err(-1, "ball %d: sigma must be > 0 and winding +-1", b);
    HIPCHK(hipStreamSynchronize(g->stream));
    if (!g->xo || !g->vo) return set_err(-1, "grid holds no emit buffers to seed");
    int rc = seed_balls_device(g->p.N, g->p.L, g->p.dx, g->kp.fs, g->mode == 2, balls, nballs, theta_sigma,
                               link_sigma, seed, g->xo, g->vo, g->stream, g->mem);
    if (rc) return set_err(-2, "device seeding failed: %s", hipGetErrorString((hipError_t)rc));
    rc = run_init(g);
    if (rc) return rc;
    g->steps = 0;
    HIPCHK(hipStreamSynchronize(g->stream));
    return 0;
}

extern "C" int scp_grid_download(scp_grid *g, double *const *cols, int ncols) {
    if (!g || !cols) return set_err(-1, "null argument");
    if (need_emit(g)) return -1;
    if (ncols != ncols_of(g->mode)) return set_err(-1, "ncols=%d, mode %d needs %d", ncols, g->mode, ncols_of(g->mode));
    const size_t nb = (size_t)g->kp.n3 * sizeof(double);
    for (int f = 0; f < g->NP; f++) {
        double *xp = cols[col_of(g->mode, f, false)], *vp = cols[col_of(g->mode, f, true)];
        if (!xp || !vp) return set_err(-1, "null column");
        HIPCHK(hipMemcpyAsync(xp, g->xo + (size_t)f * g->kp.fs, nb, hipMemcpyDeviceToHost, g->stream));
        HIPCHK(hipMemcpyAsync(vp, g->vo + (size_t)f * g->kp.fs, nb, hipMemcpyDeviceToHost, g->stream));
    }
    HIPCHK(hipStreamSynchronize(g->stream));
    return 0;
}

// nsteps launches, the last emitting (when the grid holds emit buffers); da: that
// launch also accumulates the diagnostics (march kernels of the complex and
// gauged modes only)
static int enqueue_steps(scp_grid *g, int nsteps, const DiagArgs *da = nullptr) {
    if (int rc = need_state(g)) return rc;
    if (da && need_emit(g)) return -1;
    for (int s = 0; s < nsteps; s++) {
        int rc = 0;
        if (g->aux) {
            rc = g->f32 ? kerr(launch_real_aux_f32(g->kp, g->P32[g->cur], g->W32, g->aux, g->stream), "real aux")
                        : run_aux(g, g->P[g->cur], g->W);
            if (rc) return rc;
        }
        const bool emit = (s == nsteps - 1) && g->xo;
        if (emit && da)
            rc = g->f32 ? kerr(launch_march_f32c_diag(march_mode(g), g->lo_kind, g->kp, bufs32_for(g, K_STEP), 0, *da,
                                                      g->stream), "march32 diag")
                        : kerr(launch_march_f64_diag(march_mode(g), g->kp, bufs_for(g, K_STEP), 0, *da, g->stream),
                               "march64 diag");
        else
            rc = launch_kind(g, K_STEP, emit);
        if (!rc && g->side)   // bc 1: the y edges, from this launch's side rows
            rc = kerr(launch_bc1_edges(emit, g->NP, g->kp, bufs_for(g, K_STEP), g->stream), "bc1 edges");
        if (rc) return rc;
        g->cur ^= 1;
    }
    g->steps += nsteps;
    return 0;
}

extern "C" int scp_grid_step(scp_grid *g, int nsteps) {
    if (!g) return set_err(-1, "null grid");
    if (nsteps < 1) return set_err(-1, "nsteps=%d must be >= 1", nsteps);
    int rc = enqueue_steps(g, nsteps);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(g->stream));
    return 0;
}

extern "C" int scp_grid_time_steps(scp_grid *g, int nsteps, float *ms) {
    if (!g || !ms) return set_err(-1, "null argument");
    if (nsteps < 1) return set_err(-1, "nsteps must be >= 1");
    HIPCHK(hipEventRecord(g->ev0, g->stream));
    int rc = enqueue_steps(g, nsteps);
    if (rc) return rc;
    HIPCHK(hipEventRecord(g->ev1, g->stream));
    HIPCHK(hipEventSynchronize(g->ev1));
    HIPCHK(hipEventElapsedTime(ms, g->ev0, g->ev1));
    return 0;
}

extern "C" const char *scp_grid_step_kernel_name(const scp_grid *g) { return g ? g->kname : ""; }

extern "C" int64_t scp_grid_steps_done(const scp_grid *g) { return g ? g->steps : -1; }

extern "C" int scp_grid_set_g_offset(scp_grid *g, double off) {
    if (!g) return set_err(-1, "null grid");
    g->p.G_offset = off;
    return 0;
}

extern "C" double scp_grid_g_offset(const scp_grid *g) { return g ? g->p.G_offset : 0.0; }

static int fetch(scp_grid *g, double *host, const double *dev, size_t n) {
    HIPCHK(hipMemcpyAsync(host, dev, n * sizeof(double), hipMemcpyDeviceToHost, g->stream));
    HIPCHK(hipStreamSynchronize(g->stream));
    return 0;
}

static DiagParams diag_params(const scp_grid *g) {
    const scp_params &p = g->p;
    const double dx = p.dx, dV = dx * dx * dx;
    DiagParams dp;
    memset(&dp, 0, sizeof dp);
    dp.R2core = p.qdiag_radius * p.qdiag_radius;
    const double Rint = (p.bc_type == 0) ? (p.L - p.damp_width - dx) : 1e300;   // C:2194
    dp.Rint2 = Rint * Rint;
    dp.Rcube = p.qdiag_radius;
    dp.G_offset = p.G_offset;
    dp.gauged = (g->mode == 2 && p.g != 0.0) ? 1 : 0;
    dp.has_gauge = g->mode == 2;
    dp.BPRE = dp.gauged ? 1.0 / (p.g * p.g * dx * dx * dx * dx) : 0.0;
    dp.dV = dV;
    return dp;
}

// out[] from the reduced pass-1 sums S and maxima M of the emitted state: the
// totals, then the core moments and argmax voxel (second pass) and the probes
static int diag_finish(scp_grid *g, const DiagParams &dp, const double *S, const double *M, double *out,
                       int nout) {
    for (int q = 0; q < nout; q++) out[q] = 0.0;
    const scp_params &p = g->p;
    const KParams &kp = g->kp;
    const double dx = p.dx, dV = dx * dx * dx;
    const int64_t n3 = kp.n3;
    const int nb = g->nblocks_diag;
    int rc = 0;
    out[SCP_D_E_PHI_KIN] = S[DS_EPK]; out[SCP_D_E_THETA_KIN] = S[DS_ETK];
    out[SCP_D_E_GRAD] = S[DS_EG]; out[SCP_D_E_MASS] = S[DS_EM]; out[SCP_D_E_POT] = S[DS_EP];
    out[SCP_D_E_TGRAD] = S[DS_ETG]; out[SCP_D_E_TMASS] = S[DS_ETM]; out[SCP_D_E_COUPLING] = S[DS_EC];
    double et = S[DS_EPK] + S[DS_ETK] + S[DS_EG] + S[DS_EM] + S[DS_EP] + S[DS_ETG] + S[DS_ETM] + S[DS_EC];
    if (dp.gauged) et = et + S[DS_EEM_E];
    out[SCP_D_E_TOTAL] = et;
    out[SCP_D_PHI_MAX] = M[DM_PHI];
    out[SCP_D_P_MAX] = M[DM_P];
    out[SCP_D_P_INT] = S[DS_PINT];
    out[SCP_D_THETA_RMS] = std::sqrt(S[DS_TH2] / ((g->mode == 0 ? 3.0 : 6.0) * (double)n3));
    if (g->mode == 0) return 0;

    // charges (C:2110-2168)
    out[SCP_D_Q_PHI] = S[DS_QP] * dV;
    out[SCP_D_Q_THETA] = S[DS_QT] * dV;
    out[SCP_D_Q_TOTAL] = out[SCP_D_Q_PHI] + out[SCP_D_Q_THETA];
    out[SCP_D_S_MAX] = M[DM_S];
    out[SCP_D_Q_CORE] = S[DS_QC] * dV;
    out[SCP_D_Q_P0] = S[DS_QP0] * dV; out[SCP_D_Q_P1] = S[DS_QP1] * dV; out[SCP_D_Q_P2] = S[DS_QP2] * dV;
    const int N = p.N;
    const int64_t NN = (int64_t)N * N;
    {   // theta probes (C:2142-2150)
        auto clampi = [N](long v) { return (int)(v < 0 ? 0 : (v > N - 1 ? N - 1 : v)); };
        int jc = clampi(lround(p.L / dx));
        int i1 = clampi(lround((p.qdiag_probe1 + p.L) / dx));
        int i2 = clampi(lround((p.qdiag_probe2 + p.L) / dx));
        int64_t p1 = i1 * NN + (int64_t)jc * N + jc, p2 = i2 * NN + (int64_t)jc * N + jc;
        double v[4];
        if ((rc = fetch(g, &v[0], g->xo + (size_t)(F_TU + 1) * kp.fs + p1, 1))) return rc;
        if ((rc = fetch(g, &v[1], g->xo + (size_t)(F_TV + 1) * kp.fs + p1, 1))) return rc;
        if ((rc = fetch(g, &v[2], g->xo + (size_t)(F_TU + 1) * kp.fs + p2, 1))) return rc;
        if ((rc = fetch(g, &v[3], g->xo + (size_t)(F_TV + 1) * kp.fs + p2, 1))) return rc;
        out[SCP_D_THP1U] = v[0]; out[SCP_D_THP1V] = v[1]; out[SCP_D_THP2U] = v[2]; out[SCP_D_THP2V] = v[3];
    }
    double M0 = S[DS_M0] * dV;
    double cx = 0, cy = 0, cz = 0;
    bool rcore = M0 >= 1e-30;
    if (rcore) { cx = S[DS_CX] * dV / M0; cy = S[DS_CY] * dV / M0; cz = S[DS_CZ] * dV / M0; }
    const double thr = M[DM_S] * (1.0 - 1e-12);
    unsigned long long init = ~0ull;
    HIPCHK(hipMemcpyAsync(g->argmin, &init, sizeof init, hipMemcpyHostToDevice, g->stream));
    rc = kerr(launch_diag_pass2(kp, g->xo, thr, cx, cy, cz, g->partials, g->argmin, nb, g->stream), "diag2");
    if (rc) return rc;
    rc = kerr(launch_reduce(g->partials, nb, 1, 0, g->dout, g->stream), "reduce2");
    if (rc) return rc;
    double rsum;
    if ((rc = fetch(g, &rsum, g->dout, 1))) return rc;
    unsigned long long am;
    HIPCHK(hipMemcpyAsync(&am, g->argmin, sizeof am, hipMemcpyDeviceToHost, g->stream));
    HIPCHK(hipStreamSynchronize(g->stream));
    out[SCP_D_R_CORE] = rcore ? std::sqrt(rsum * dV / M0) : 0.0;
    out[SCP_D_OMEGA_CORE] = 0.0;
    if (am != ~0ull) {
        double u0, v0, ud0, vd0;
        if ((rc = fetch(g, &u0, g->xo + (size_t)F_U * kp.fs + am, 1))) return rc;
        if ((rc = fetch(g, &v0, g->xo + (size_t)F_V * kp.fs + am, 1))) return rc;
        if ((rc = fetch(g, &ud0, g->vo + (size_t)F_U * kp.fs + am, 1))) return rc;
        if ((rc = fetch(g, &vd0, g->vo + (size_t)F_V * kp.fs + am, 1))) return rc;
        double den = u0 * u0 + v0 * v0;
        if (den >= 1e-20) out[SCP_D_OMEGA_CORE] = (u0 * vd0 - v0 * ud0) / den;
    }
    if (g->mode == 2) {   // compute_gauss (C:2239-2242)
        out[SCP_D_GAUSS_MAX] = M[DM_GAUSS];
        out[SCP_D_GAUSS_L2] = S[DS_NOMEGA] > 0 ? std::sqrt(S[DS_GL2] / S[DS_NOMEGA]) : 0.0;
        out[SCP_D_E_EM] = S[DS_EEM_G];
        out[SCP_D_Q_FLUX] = (p.g != 0) ? (S[DS_FSUM] - p.G_offset * S[DS_NC] * dV) / p.g : 0.0;
    }
    return 0;
}

extern "C" int scp_grid_diag(scp_grid *g, double *out, int nout) {
    if (!g || !out) return set_err(-1, "null argument");
    if (need_emit(g)) return -1;
    if (nout < SCP_D_COUNT) return set_err(-1, "nout=%d < %d", nout, SCP_D_COUNT);
    const DiagParams dp = diag_params(g);
    const int nb = g->nblocks_diag;
    int rc = kerr(launch_diag_pass1(g->mode, g->kp, dp, g->xo, g->vo, g->partials, nb, g->stream), "diag1");
    if (rc) return rc;
    rc = kerr(launch_reduce(g->partials, nb, DS_NSUM, DM_NMAX, g->dout, g->stream), "reduce");
    if (rc) return rc;
    double r[DS_NSUM + DM_NMAX];
    if ((rc = fetch(g, r, g->dout, DS_NSUM + DM_NMAX))) return rc;
    return diag_finish(g, dp, r, r + DS_NSUM, out, nout);
}

// Steps with the diagnostics of the final state fused into the emitting
// launch: the march accumulates every pass-1 sum and maximum of the emitted
// state per block (and the gauged charge density), a light pass adds the
// Gauss-law statistics, which need E at neighbours other blocks compute. Grids
// whose step does not run the march (direct-neighbour kernel, gauged storage
// with g = 0) take the step and the stand-alone pass.

// the fused path's work space: partial rows (nbm march blocks + nbt Gauss-tail
// blocks, DS_NSUM + DM_NMAX doubles each) and, gauged, N^3 doubles of charge density
static bool fused_diag(const scp_grid *g) { return use_march(g, K_STEP) && g->step_mode == g->mode; }
static void fused_rows(const scp_grid *g, int *nbm, int *nbt) {
    *nbm = march_blocks(g->kp.N, 0);
    *nbt = g->mode == 2 ? g->nblocks_diag : 0;
}

// the fused step + diagnostics over caller-provided work space (fpart: the
// partial rows, every row written by its block before it is read; rho: N^3,
// gauged, every site written by the march before the Gauss tail reads it)
static int step_diag_run(scp_grid *g, int nsteps, double *out, int nout, double *fpart, double *rho) {
    constexpr int NQ = DS_NSUM + DM_NMAX;
    int nbm, nbt;
    fused_rows(g, &nbm, &nbt);
    DiagArgs da;
    da.dp = diag_params(g);
    da.rho = rho;
    da.partials = fpart;
    int rc = enqueue_steps(g, nsteps, &da);
    if (rc) return rc;
    if (nbt)
        rc = kerr(launch_gauss_seam(g->kp, da.dp, 0, g->vo, rho, fpart + (size_t)nbm * NQ, nbt, g->stream),
                  "gauss seam");
    if (rc) return rc;
    rc = kerr(launch_reduce_wide(fpart, nbm + nbt, DS_NSUM, DM_NMAX, g->dout, g->stream), "reduce");
    if (rc) return rc;
    double r[NQ];
    if ((rc = fetch(g, r, g->dout, NQ))) return rc;
    return diag_finish(g, da.dp, r, r + DS_NSUM, out, nout);
}

extern "C" int scp_grid_step_diag(scp_grid *g, int nsteps, double *out, int nout) {
    if (!g || !out) return set_err(-1, "null argument");
    if (nsteps < 1) return set_err(-1, "nsteps=%d must be >= 1", nsteps);
    if (nout < SCP_D_COUNT) return set_err(-1, "nout=%d < %d", nout, SCP_D_COUNT);
    if (!fused_diag(g)) {
        int rc = scp_grid_step(g, nsteps);
        return rc ? rc : scp_grid_diag(g, out, nout);
    }
    constexpr int NQ = DS_NSUM + DM_NMAX;
    int nbm, nbt;
    fused_rows(g, &nbm, &nbt);
    if (g->fpart_rows < nbm + nbt) {
        grelease(g, g->fpart);
        g->fpart_rows = 0;
        HIPCHK(galloc(g, &g->fpart, (size_t)(nbm + nbt) * NQ * sizeof(double)));
        g->fpart_rows = nbm + nbt;
    }
    if (g->mode == 2 && !g->rho) HIPCHK(galloc(g, &g->rho, (size_t)g->kp.n3 * sizeof(double)));
    return step_diag_run(g, nsteps, out, nout, g->fpart, g->rho);
}

// ---------------------------------------------------------------------------
// The allocation-free variant: the same work space in a static device array.
// Each device the library runs on holds one copy (SCRATCH_BYTES) until the
// process ends; its address is looked up once per device, with that device
// current, and only that address is used from the host.
// ---------------------------------------------------------------------------
constexpr size_t SCRATCH_BYTES = (size_t)64 << 20;
constexpr int SCRATCH_MAX_DEVICES = 64;
static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];
static std::mutex scratch_mu;
static void *scratch_dev[SCRATCH_MAX_DEVICES];

static int scratch_for(int dev, void **ptr) {
    if (dev < 0 || dev >= SCRATCH_MAX_DEVICES)
        return set_err(-1, "device %d: the scratch table holds devices 0..%d", dev, SCRATCH_MAX_DEVICES - 1);
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_dev[dev]) {
        int cur = -1;
        HIPCHK(hipGetDevice(&cur));
        if (cur != dev) HIPCHK(hipSetDevice(dev));
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != dev) HIPCHK(hipSetDevice(cur));
        if (e != hipSuccess || !p)
            return set_err(-2, "hipGetSymbolAddress(scratch) on device %d failed: %s", dev, hipGetErrorString(e));
        scratch_dev[dev] = p;
    }
    *ptr = scratch_dev[dev];
    return 0;
}

extern "C" int scp_grid_step_diag_noalloc(scp_grid *g, int nsteps, double *out, int nout) {
    if (!g || !out) return set_err(-1, "null argument");
    if (nsteps < 1) return set_err(-1, "nsteps=%d must be >= 1", nsteps);
    if (nout < SCP_D_COUNT) return set_err(-1, "nout=%d < %d", nout, SCP_D_COUNT);
    if (!fused_diag(g)) {   // step and stand-alone pass: grid-owned buffers only
        int rc = scp_grid_step(g, nsteps);
        return rc ? rc : scp_grid_diag(g, out, nout);
    }
    constexpr int NQ = DS_NSUM + DM_NMAX;
    int nbm, nbt;
    fused_rows(g, &nbm, &nbt);
    const size_t rows = ((size_t)(nbm + nbt) * NQ * sizeof(double) + 255) / 256 * 256;
    const size_t need = rows + (g->mode == 2 ? (size_t)g->kp.n3 * sizeof(double) : 0);
    if (need > SCRATCH_BYTES)
        return set_err(-1, "N=%d needs %zu bytes of diagnostics work space, above the %zu of the static "
                           "scratch (gauged grids up to N^3 ~ 8.3e6 sites, N <= 202)",
                       g->p.N, need, SCRATCH_BYTES);
    void *base = nullptr;
    int rc = scratch_for(g->device, &base);
    if (rc) return rc;
    double *fpart = (double *)base;
    double *rho = g->mode == 2 ? (double *)((char *)base + rows) : nullptr;
    return step_diag_run(g, nsteps, out, nout, fpart, rho);
}

extern "C" int scp_grid_forces(scp_grid *g, double *const *acc, int nacc) {
    if (!g || !acc) return set_err(-1, "null argument");
    if (need_emit(g)) return -1;
    if (need_state(g)) return -1;
    const int want = g->mode == 2 ? 15 : (g->mode == 1 ? 12 : 6);
    if (nacc != want) return set_err(-1, "nacc=%d, mode %d needs %d", nacc, g->mode, want);
    if (!g->acc) HIPCHK(galloc(g, &g->acc, (size_t)g->NP * g->kp.fs * sizeof(double)));
    HIPCHK(hipMemsetAsync(g->acc, 0, (size_t)g->NP * g->kp.fs * sizeof(double), g->stream));
    int rc = run_aux(g, g->xo, g->vo);
    if (rc) return rc;
    rc = launch_kind(g, K_FORCES, false);
    if (rc) return rc;
    const size_t nb = (size_t)g->kp.n3 * sizeof(double);
    for (int f = 0; f < g->NP; f++)
        HIPCHK(hipMemcpyAsync(acc[acc_col_of(g->mode, f)], g->acc + (size_t)f * g->kp.fs, nb,
                              hipMemcpyDeviceToHost, g->stream));
    HIPCHK(hipStreamSynchronize(g->stream));
    return 0;
}

extern "C" int scp_grid_vec_fit(scp_grid *g, int block_size, float *coeffs, int64_t n_coeffs) {
    if (!g || !coeffs) return set_err(-1, "null argument");
    if (need_emit(g)) return -1;
    if (g->mode != 0)
        return set_err(-1, "vector frames are real-mode output (complex_phi=1 refuses vec_snap_dt, "
                           "scp_config.h:458)");
    if (block_size < 4 || block_size > 16 || block_size > g->kp.N)
        return set_err(-1, "vec_block_size=%d outside [4, min(16, N)] (a tricubic fit needs 4 samples per axis)",
                       block_size);
    const int64_t bn = g->kp.N / block_size, want = bn * bn * bn * 384;
    if (n_coeffs != want) return set_err(-1, "n_coeffs=%lld, (N/bs)^3 x 384 = %lld", (long long)n_coeffs, (long long)want);
    float *d = nullptr;
    HIPCHK(hipMalloc(&d, (size_t)want * sizeof(float)));
    int rc = kerr(launch_vec_fit(g->xo, g->kp.fs, g->kp.N, block_size, d, g->stream), "vec fit");
    hipError_t e = rc ? hipSuccess : hipMemcpyAsync(coeffs, d, (size_t)want * sizeof(float), hipMemcpyDeviceToHost, g->stream);
    if (!rc && e == hipSuccess) e = hipStreamSynchronize(g->stream);
    hipFree(d);
    if (rc) return rc;
    HIPCHK(e);
    return 0;
}

extern "C" int scp_grid_rescale_phi(scp_grid *g, double f) {
    if (!g) return set_err(-1, "null grid");
    if (need_emit(g)) return -1;
    if (need_state(g)) return -1;
    // Between steps the reference scales phi and its velocity in place and keeps
    // the stored acceleration a_n = F(x_n) of the unscaled state: the next half
    // kick uses it, the next force evaluation sees the scaled positions
    // (rescale_phi_kernel, scp_sim.cu:2619-2624, called at scp_sim.cu:5512-5517).
    // a_n is recomputed from x_n before the scaling (the forces kernel is the
    // step's own, so it is the step's a_n; with sigma_freq the theta frequency
    // is then taken from the completed velocities).
    const size_t nacc = (size_t)g->NP * g->kp.fs * sizeof(double);
    if (!g->acc) HIPCHK(galloc(g, &g->acc, nacc));
    int rc = run_aux(g, g->xo, g->vo);
    if (!rc) rc = launch_kind(g, K_FORCES, false);
    if (rc) return rc;
    const int nphi = g->mode == 0 ? 3 : 6;    // u (and v)
    for (int q = 0; q < nphi; q++) {
        int fld = q < 3 ? F_U + q : F_V + (q - 3);
        rc = kerr(launch_scale(g->xo + (size_t)fld * g->kp.fs, f, g->kp.n3, g->stream), "scale");
        if (rc) return rc;
        rc = kerr(launch_scale(g->vo + (size_t)fld * g->kp.fs, f, g->kp.n3, g->stream), "scale");
        if (rc) return rc;
    }
    g->kp.init_acc = 1;
    rc = run_init(g, false);
    g->kp.init_acc = 0;
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(g->stream));
    return 0;
}

extern "C" int scp_grid_set_kappa(scp_grid *g, double kappa) {
    if (!g) return set_err(-1, "null grid");
    if (!(kappa > 0)) return set_err(-1, "kappa=%g must be > 0", kappa);
    g->kp.KAPPA = kappa;
    g->p.kappa = kappa;
    return 0;
}

extern "C" int scp_grid_set_damp_rate(scp_grid *g, double rate) {
    if (!g) return set_err(-1, "null grid");
    if (!(rate >= 0)) return set_err(-1, "damp_rate=%g must be >= 0", rate);
    g->p.damp_rate = rate;
    g->kp.DR = rate;
    // the sponge stage runs where make_kparams enables it: bc 0 with a positive width and rate
    g->kp.damp = (g->p.bc_type == 0 && g->p.damp_width > 0 && rate > 0) ? 1 : 0;
    return 0;
}

// Gauss projection on the device, same algorithm and stopping rule as the
// reference (CG on -lap chi = G - Gbar, chi kept zero-mean, tol 1e-13*max(1,|b|max),
// at most 20000 iterations), in FP64.
extern "C" int scp_grid_gauss_project(scp_grid *g, int *iters, double *gauss_max) {
    if (!g) return set_err(-1, "null grid");
    if (need_emit(g)) return -1;
    if (need_state(g)) return -1;
    if (g->mode != 2) return set_err(-1, "gauss projection needs the gauged mode");
    const KParams &kp = g->kp;
    const uint32_t n = kp.n3;
    const double N3 = (double)n;
    const int nb = g->nblocks_diag;
    double *b = nullptr, *x = nullptr, *r = nullptr, *q = nullptr, *Aq = nullptr;
    const size_t bytes = (size_t)n * sizeof(double);
    int rc = 0;
    auto cleanup = [&]() { gfree(g, b); gfree(g, x); gfree(g, r); gfree(g, q); gfree(g, Aq); };
    if (galloc(g, &b, bytes) || galloc(g, &x, bytes) || galloc(g, &r, bytes) || galloc(g, &q, bytes) ||
        galloc(g, &Aq, bytes)) {
        cleanup();
        return set_err(-2, "gauss projection alloc failed");
    }
    double hv[3];
    // b = G; gbar; b -= gbar; gmax0
    rc = kerr(launch_gauss_residual(kp, g->xo, g->vo, b, g->stream), "gauss residual");
    // sum of b: dot with a ones vector stored in x temporarily
    if (!rc) {
        std::vector<double> ones(n, 1.0);
        if (hipMemcpyAsync(x, ones.data(), bytes, hipMemcpyHostToDevice, g->stream)) rc = set_err(-2, "copy");
        if (!rc) rc = kerr(launch_dot_partials(b, x, n, g->partials, nb, g->stream), "dot");
        if (!rc) rc = kerr(launch_reduce(g->partials, nb, 1, 0, g->dout, g->stream), "reduce");
        if (!rc) rc = fetch(g, hv, g->dout, 1);
        hipStreamSynchronize(g->stream);
    }
    double gbar = rc ? 0 : hv[0] / N3, gmax0 = 0;
    if (!rc) rc = kerr(launch_sub_mean_max(b, gbar, n, g->partials, nb, g->stream), "submean");
    if (!rc) rc = kerr(launch_reduce(g->partials, nb, 0, 1, g->dout, g->stream), "reduce");
    if (!rc) rc = fetch(g, &gmax0, g->dout, 1);
    const double tol = 1e-13 * std::fmax(1.0, gmax0);
    if (!rc && (hipMemsetAsync(x, 0, bytes, g->stream) || hipMemcpyAsync(r, b, bytes, hipMemcpyDeviceToDevice, g->stream) ||
                hipMemcpyAsync(q, b, bytes, hipMemcpyDeviceToDevice, g->stream)))
        rc = set_err(-2, "copy");
    double rz = 0;
    if (!rc) rc = kerr(launch_dot_partials(r, r, n, g->partials, nb, g->stream), "dot");
    if (!rc) rc = kerr(launch_reduce(g->partials, nb, 1, 0, g->dout, g->stream), "reduce");
    if (!rc) rc = fetch(g, &rz, g->dout, 1);
    int it = 0;
    double rmax = gmax0;
    while (!rc && rmax > tol && it < 20000) {
        double qAq = 0;
        rc = kerr(launch_neg_lap(kp, q, Aq, g->stream), "neglap");
        if (!rc) rc = kerr(launch_dot_partials(q, Aq, n, g->partials, nb, g->stream), "dot");
        if (!rc) rc = kerr(launch_reduce(g->partials, nb, 1, 0, g->dout, g->stream), "reduce");
        if (!rc) rc = fetch(g, &qAq, g->dout, 1);
        if (rc) break;
        double alpha = rz / qAq;
        rc = kerr(launch_axpy_cg(alpha, q, Aq, x, r, n, g->partials, nb, g->stream), "axpy");
        if (!rc) rc = kerr(launch_reduce(g->partials, nb, 2, 1, g->dout, g->stream), "reduce");
        if (!rc) rc = fetch(g, hv, g->dout, 3);
        if (rc) break;
        double rz2 = hv[0], xbar = hv[1] / N3, rm = hv[2];
        double beta = rz2 / rz;
        rc = kerr(launch_cg_update(xbar, beta, r, x, q, n, g->stream), "cgupdate");
        rz = rz2;
        rmax = rm;
        it++;
    }
    if (!rc && rmax > tol)
        rc = set_err(-5, "Gauss projection CG failed to converge (%d iters, max|res|=%.3e > tol=%.3e)", it, rmax, tol);
    if (!rc) rc = kerr(launch_add_grad(kp, x, g->vo + (size_t)F_TH * kp.fs, g->stream), "addgrad");
    // G_offset = mean of the corrected residual; report max |G - G_offset|
    if (!rc) rc = kerr(launch_gauss_residual(kp, g->xo, g->vo, b, g->stream), "gauss residual");
    if (!rc) {
        std::vector<double> ones(n, 1.0);
        if (hipMemcpyAsync(x, ones.data(), bytes, hipMemcpyHostToDevice, g->stream)) rc = set_err(-2, "copy");
        if (!rc) rc = kerr(launch_dot_partials(b, x, n, g->partials, nb, g->stream), "dot");
        if (!rc) rc = kerr(launch_reduce(g->partials, nb, 1, 0, g->dout, g->stream), "reduce");
        if (!rc) rc = fetch(g, hv, g->dout, 1);
        hipStreamSynchronize(g->stream);
    }
    double gres = 0;
    if (!rc) {
        g->p.G_offset = hv[0] / N3;
        rc = kerr(launch_sub_mean_max(b, g->p.G_offset, n, g->partials, nb, g->stream), "submean");
        if (!rc) rc = kerr(launch_reduce(g->partials, nb, 0, 1, g->dout, g->stream), "reduce");
        if (!rc) rc = fetch(g, &gres, g->dout, 1);
    }
    if (!rc) rc = run_init(g);
    hipStreamSynchronize(g->stream);
    cleanup();
    if (rc) return rc;
    if (iters) *iters = it;
    if (gauss_max) *gauss_max = gres;
    return 0;
}
