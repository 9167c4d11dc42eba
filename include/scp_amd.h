/* scp_amd.h — C-ABI of the MI355X sfa/sim leapfrog step (libscp_amd.so).
 *
 * This is the boundary a host program binds (the drop-in scp_sim binary in
 * scp_amd/csrc/host, Python via ctypes, Go via cgo — see INTEGRATION.md). Plain
 * C types only: pointers, sizes, int status codes (0 = ok, <0 = error, message
 * via scp_last_error). No process exit inside the library.
 *
 * Each entry point replaces one reference interface:
 *   scp_grid_create    grid_alloc (sfa/sim/scp_sim.c:199-367) + gpu_alloc /
 *                      gpu_set_constants (sfa/sim/scp_sim.cu:2659-2707, 2915-2948)
 *   scp_grid_upload    gpu_upload (scp_sim.cu:2768-2820) + the initial force
 *                      call (scp_sim.c:2833-2838)
 *   scp_grid_step      n x verlet_step (scp_sim.c:1583-1716) / gpu_verlet_step
 *                      (scp_sim.cu:2971-3208), boundary stage included
 *   scp_grid_download  gpu_download (scp_sim.cu:2822-2864); the 12/24/30 snapshot
 *                      columns of sfa_snap (scp_sim.c:2646-2700)
 *   scp_grid_diag      compute_energy* + P_integrated* + theta_rms* +
 *                      compute_charges + compute_gauss (scp_sim.c:1722-2243) /
 *                      run_gpu_diagnostics (scp_sim.cu:3697-4041)
 *   scp_grid_forces    compute_forces / _complex / _complex_gauge
 *                      (scp_sim.c:743-994, 1053-1263): accelerations of the
 *                      current state (used by the gauge-transform self-test,
 *                      scp_sim.c:2461-2556)
 *   scp_grid_gauss_project  init_gauss_project (scp_sim.c:2284-2400) on the GPU
 *   scp_grid_rescale_phi    rescale_phi_kernel (scp_sim.cu:2619-2624)
 *   scp_grid_set_kappa      the self-tune kappa update (scp_sim.cu:5520-5541)
 *   scp_grid_set_damp_rate  the absorbing -> periodic switch at bc_switch_time
 *                           (scp_sim.c:3073-3081, scp_sim.cu:5464-5490)
 *   scp_grid_vec_fit   fit_patch over the six fields of a vector frame
 *                      (scp_sim.c:2567-2640, 3106-3117)
 *   scp_qflow_*        the fixed-charge gradient flow of the seed relaxer
 *                      sfa/seed/eta_qflow.c:242-375 (fixed-Q mode)
 *
 * Column order of every host state buffer = the reference snapshot order
 * (scp_sim.c:2862-2897): phi_x..z, theta_x..z, phi_vx..vz, theta_vx..vz,
 * [phiim, thetaim, phiim_v, thetaim_v], [th_x..z, E_x..z]; site index
 * i*N*N + j*N + k (z fastest).
 */
#ifndef SCP_AMD_H
#define SCP_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SCP_ABI_VERSION 1

enum { SCP_MODE_REAL = 0, SCP_MODE_COMPLEX = 1, SCP_MODE_GAUGED = 2 };
enum { SCP_PREC_F64 = 0, SCP_PREC_F32C = 1 };

typedef struct {
    int32_t N;            /* sites per edge (periodic cube) */
    int32_t mode;         /* SCP_MODE_* ; gauged with g == 0 runs the complex step (scp_sim.c:1586) */
    int32_t precision;    /* SCP_PREC_F64: FP64 state and stencil (parity path);
                             SCP_PREC_F32C: FP32 state + FP32 stencil with compensated
                             (two-sum) position/velocity accumulation */
    int32_t bc_type;      /* 0 absorbing sphere, 1 gradient pinned (real and complex modes,
                             FP64), 2 periodic */
    int32_t rmode;        /* real-mode mass coupling: 0 const, 1 inverse, 3 density-kappa */
    int32_t lo_words;     /* SCP_PREC_F32C compensation-word storage: 0 FP32 (8 B/scalar),
                             1 bfloat16 (6 B/scalar), 2 none (plain FP32, 4 B/scalar),
                             3 signed byte in units of ulp(hi)/256 (5 B/scalar),
                             4 such bytes on the positions only (4.5 B/scalar) */
    int32_t step_kernel;  /* 0: LDS-staged march (default, every mode; bc_type 1 and one
                                field above 4 GiB run the direct kernel), 1: direct-neighbour
                                march */
    int32_t gradient_margin;  /* bc_type 1: x-slabs pinned on each face (gradient_margin,
                                 scp_config.h:110); must be 0 otherwise */
    double L, dx, dt;     /* dx = 2L/(N-1), dt = dt_factor*dx (scp_sim.c:204-205) */
    double m2, mtheta2, eta, mu, kappa, g;
    double kappa_h, alpha_cs, beta_h, inv_alpha, inv_beta, kappa_gamma;
    double damp_width, damp_rate;
    double G_offset;      /* frozen jellium Gauss offset (set by gauss_project or the host) */
    double qdiag_radius, qdiag_probe1, qdiag_probe2;
    /* Real mode only: the terms the reference applies in its GPU step and not
     * in its CPU step (scp_sim.cu:1144-1248), 0 = off: theta saturation and
     * conversion (theta_sat, gamma_conv), gradient self-interaction with
     * chirality bias (sigma_grad, chi_chiral), cubic self-interaction,
     * frequency-mismatch conversion, cross potential, theta self potential. */
    double theta_sat, gamma_conv, sigma_grad, chi_chiral, sigma_cubic, sigma_freq, sigma_cross, lambda_self;
} scp_params;

/* Diagnostics vector filled by scp_grid_diag: every value of one diag.tsv row
 * (scp_sim.c:3040-3065), FP64, in this order. */
enum {
    SCP_D_E_PHI_KIN = 0, SCP_D_E_THETA_KIN, SCP_D_E_GRAD, SCP_D_E_MASS, SCP_D_E_POT,
    SCP_D_E_TGRAD, SCP_D_E_TMASS, SCP_D_E_COUPLING, SCP_D_E_TOTAL, SCP_D_PHI_MAX,
    SCP_D_P_MAX, SCP_D_P_INT, SCP_D_THETA_RMS,
    SCP_D_Q_PHI, SCP_D_Q_THETA, SCP_D_Q_TOTAL, SCP_D_S_MAX, SCP_D_R_CORE, SCP_D_OMEGA_CORE,
    SCP_D_Q_CORE, SCP_D_THP1U, SCP_D_THP1V, SCP_D_THP2U, SCP_D_THP2V,
    SCP_D_GAUSS_MAX, SCP_D_GAUSS_L2, SCP_D_E_EM, SCP_D_Q_FLUX,
    SCP_D_Q_P0, SCP_D_Q_P1, SCP_D_Q_P2,
    SCP_D_COUNT
};

typedef struct scp_grid scp_grid;

int  scp_abi_version(void);
/* Copies the last error message of the calling thread; returns its length. */
int  scp_last_error(char *buf, int buflen);
/* Device HBM bytes a grid with these parameters allocates (admission control). */
int64_t scp_grid_bytes(const scp_params *p);

int  scp_grid_create(const scp_params *p, scp_grid **out);
int  scp_grid_destroy(scp_grid *g);
/* ncols must be 12/24/30 for real/complex/gauged. Uploads the state, runs the
 * initial force evaluation and the first half-kick + drift. */
int  scp_grid_upload(scp_grid *g, const double *const *cols, int ncols);
/* Current positions and velocities (state after the last completed step). */
int  scp_grid_download(scp_grid *g, double *const *cols, int ncols);
/* Advance nsteps leapfrog steps (nsteps >= 1). */
int  scp_grid_step(scp_grid *g, int nsteps);
int  scp_grid_diag(scp_grid *g, double *out, int nout);
/* scp_grid_step followed by scp_grid_diag of the resulting state, with the
 * diagnostics fused into the step's last launch (grids on the LDS-staged
 * march: every energy, charge and moment sum is reduced by the step kernel
 * itself; gauged, so are the Gauss-law statistics of every site whose
 * backward neighbours the same block emits, and a light pass adds those of
 * the tile and segment seams).
 * Replaces the main loop's verlet_step + diagnostics pair on a diag step
 * (scp_sim.c:3082 and 2993-3065; scp_sim.cu:5492, 3697-4041). Other grids
 * run the step and the stand-alone pass. The first call on a gauged grid
 * allocates N^3 doubles for the seam sites' charge density. */
int  scp_grid_step_diag(scp_grid *g, int nsteps, double *out, int nout);
/* scp_grid_step_diag without any device allocation. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The work space of the fused diagnostics (per-block partial rows and, gauged,
 * N^3 doubles of charge density) lives in a static 64 MiB device array that
 * every GPU the library runs on holds until the process ends; grids needing
 * more (gauged N > ~202) are rejected with an error. Results equal
 * scp_grid_step_diag's. Every call on one GPU shares that GPU's array, so
 * calls that overlap on one GPU (different host threads) race: serialise them. */
int  scp_grid_step_diag_noalloc(scp_grid *g, int nsteps, double *out, int nout);
/* Accelerations of the current positions: 6/12/15 columns in the order
 * phi_acc, theta_acc, [phi_im_acc, theta_im_acc], [E_acc]. */
int  scp_grid_forces(scp_grid *g, double *const *acc, int nacc);
/* Vector-frame fits of the real mode: the tricubic Chebyshev least-squares
 * coefficients of every block_size^3 patch (patch p at origin
 * (bi, bj, bk) * block_size, p = (bi * bn + bj) * bn + bk, bn = N / block_size)
 * of phi_x..z and theta_x..z (the current state rounded to FP32), 384 floats
 * per patch: field f at 64 f + 16 a + 4 b + c for T_a(x) T_b(y) T_c(z). Replaces
 * fit_patch (scp_sim.c:2567-2640) over the six fields (scp_sim.c:3106-3117);
 * 4 <= block_size <= 16. n_coeffs must be bn^3 * 384. The temporal model and
 * the FRVD frames stay on the host (libscp_io: vecframes.h). */
int  scp_grid_vec_fit(scp_grid *g, int block_size, float *coeffs, int64_t n_coeffs);
/* Gauss projection of the current E (gauged mode): CG solve, E correction,
 * G_offset freeze. Writes iterations and the post-projection max|G - G_offset|. */
int  scp_grid_gauss_project(scp_grid *g, int *iters, double *gauss_max);
int  scp_grid_set_g_offset(scp_grid *g, double g_offset);
double scp_grid_g_offset(const scp_grid *g);
/* phi, phi_vel (and imaginary parts) *= f between steps — the fixed-charge
 * projection of the self-tune loop (rescale_phi_kernel, scp_sim.cu:2619-2624,
 * scp_sim.cu:5512-5517). As there, the next half kick uses the acceleration of
 * the state before the scaling; the next force evaluation sees the scaled one. */
int  scp_grid_rescale_phi(scp_grid *g, double f);
/* kappa of the potential for the following steps — the self-tune update
 * (cudaMemcpyToSymbol(d_KAPPA), scp_sim.cu:5531-5541); kappa > 0. */
int  scp_grid_set_kappa(scp_grid *g, double kappa);
/* damp_rate of the absorbing sphere (bc_type 0) for the following steps; rate
 * >= 0, 0 turns the sponge off. The reference's bc_switch_time sets damp_rate
 * = 0 before the step lround(bc_switch_time / dt) (c.damp_rate = 0.0,
 * scp_sim.c:3077-3081; cudaMemcpyToSymbol(d_DAMP_RATE), scp_sim.cu:5485-5490):
 * a caller switching there calls this after the previous step completed. */
int  scp_grid_set_damp_rate(scp_grid *g, double rate);
/* Steps completed since upload. */
int64_t scp_grid_steps_done(const scp_grid *g);

/* Timing of the step kernel on the grid's own stream (HIP events around
 * nsteps launches; no host synchronisation inside the timed region). */
int  scp_grid_time_steps(scp_grid *g, int nsteps, float *ms_total);
/* Name of the step kernel as rocprof reports it (for profile matching). */
const char *scp_grid_step_kernel_name(const scp_grid *g);

/* One Gaussian Q-ball of an initial state (physical units, minimum image in
 * the periodic box): Phi_a = amp exp(-r^2 / (2 sigma^2)) e^{i(phase + winding
 * 2 pi a / 3)}, dPhi_a/dt = i omega Phi_a. */
typedef struct {
    double x0, y0, z0, amp, sigma, omega, phase;
    int32_t winding;      /* +1 / -1 */
    int32_t reserved;     /* 0 */
} scp_qball;

/* Seed the grid on the device (complex or gauged mode): a superposition of
 * Q-balls, Theta ~ N(0, theta_sigma^2) per component (tu and tv), zero Theta
 * velocities, links ~ N(0, link_sigma^2), E = 0, from counter-based random
 * numbers keyed by seed; then the initial force evaluation and first half kick,
 * as scp_grid_upload. The reference seeds such states with separate generator
 * programs that write an SFA file for init=sfa (scp_init.h:489-499); this
 * needs neither the host state nor its copy. */
int  scp_grid_seed_balls(scp_grid *g, const scp_qball *balls, int nballs, double theta_sigma,
                         double link_sigma, uint64_t seed);

/* Batches: many grids advanced together. No single-call counterpart in the
 * reference, where every job is its own scp_sim process started by the Go
 * runner (LocalExecutor.executeRun, sfa/runner/local.go:239-330) with at most
 * max(1, nproc/16) running at once (localMaxConcurrent, local.go:51-64). A
 * batch holds grids of one class (kernel mode, precision, lo_words) that run
 * the LDS-staged march; one launch per step covers every member's tiles, and
 * chains of steps replay captured HIP graphs. Grids stay owned by the caller;
 * a grid belongs to at most one batch and leaves it when destroyed. Results
 * are bitwise those of scp_grid_step on each grid alone. */
typedef struct scp_batch scp_batch;

int  scp_batch_create(scp_batch **out);
int  scp_batch_destroy(scp_batch *b);
int  scp_batch_add(scp_batch *b, scp_grid *g);
int  scp_batch_remove(scp_batch *b, scp_grid *g);
int  scp_batch_size(const scp_batch *b);
/* Member i (order of addition; removals close the gaps) advances nsteps[i] >= 0
 * steps; n must equal scp_batch_size. */
int  scp_batch_step(scp_batch *b, const int *nsteps, int n);
int  scp_batch_step_all(scp_batch *b, int nsteps);
/* scp_batch_step_all timed with HIP events around the graph replay. */
int  scp_batch_time_steps(scp_batch *b, int nsteps, float *ms_total);
/* Autotuner: tile shape (64 x 4 sites; for FP32C batches with int8 position
 * words also 128 x 4, 64 x 8, 64 x 2 and 32 x 8) x planes per block (default, 16, 32,
 * 64, 128) x launch form (one block per tile, or persistent resident blocks
 * over contiguous tile ranges), each timed on a 3-step chain of the batch's
 * own steps; every member advances *steps_taken steps (results do not depend
 * on the configuration). */
int  scp_batch_tune(scp_batch *b, int *steps_taken);
/* Tile of the batched launch: tk (k) x tj (j) sites, (tk, tj) in (64, 4),
 * (128, 4), (64, 8), (64, 2), (32, 8), or (0, 0): each member its own of 64 x 4
 * and 32 x 8 in one launch; all but (64, 4) for FP32C batches with int8
 * position words only. The 32 x 8 tile runs rows of 32 sites, two per wave:
 * grids whose edge is not a multiple of 64 leave fewer lanes idle. Until this
 * call or the autotuner picks one, such a batch gives each member the shape it
 * runs fastest (32 x 8 where it issues at most 90 % of the lanes of 64 x 4, or
 * no more lanes at edges <= 128), re-decided as members change (results do not
 * depend on the shape); every other batch runs 64 x 4. scp_batch_tile reports
 * (0, 0) for the per-member launch. */
int  scp_batch_set_tile(scp_batch *b, int tk, int tj);
int  scp_batch_tile(const scp_batch *b, int *tk, int *tj);
/* di = 0: default depth (the deepest of 128, 64, 32, 16 planes that gives
 * >= 16384 blocks over all members); persistent: 0 / 1. */
int  scp_batch_set_config(scp_batch *b, int di, int persistent);
int  scp_batch_config(const scp_batch *b, int *di, int *persistent, int *tuned);

/* Serving: a request loop over device-resident grids. Counterpart of the Go
 * runner's job model (sfa/runner/local.go: Run :213 queues a config, a
 * concurrency slot is acquired :253, executeRun :239 execs one scp_sim process
 * per job, RunStatus :483, Download :514), rebuilt as one process whose jobs
 * share the GPU: jobs are (parameters, Q-balls, seed, steps); each round
 * admits queued jobs in arrival order while their device bytes (scp_grid_bytes
 * plus the projection's work arrays) fit the HBM budget (seeded on the device,
 * Gauss-projected if asked), advances all active jobs by up to `quantum` steps
 * in one batched chain per class, and retires finished jobs, whose grids stay
 * resident until scp_server_take. A queued job that does not fit is passed
 * over by later jobs that do for at most `aging` rounds (scp_server_set_aging,
 * default 8; 0 = strict arrival order). Device memory comes from a pool the
 * server grows in slabs of >= 2 GiB and reuses; a job holds its FP64 emit
 * buffers only at admission and from its last chain on, its step state only
 * until its last step. Take finished results: they hold memory until then. */
typedef struct scp_server scp_server;

typedef struct {
    scp_params params;         /* complex or gauged; step_kernel 0 */
    const scp_qball *balls;    /* copied at submission */
    int32_t nballs;
    int32_t nsteps;            /* >= 1 */
    double theta_sigma, link_sigma;
    uint64_t seed;
    int32_t gauss_project;     /* 1: Gauss projection after seeding (gauged, g != 0) */
    int32_t reserved;          /* 0 */
} scp_job;

enum { SCP_JOB_FAILED = -1, SCP_JOB_QUEUED = 0, SCP_JOB_ACTIVE = 1, SCP_JOB_DONE = 2 };

typedef struct {
    int32_t state;             /* SCP_JOB_* */
    int32_t nsteps;
    int64_t steps_done, bytes; /* bytes: device memory the job holds now */
    double t_submit, t_admit, t_done;   /* seconds on the server's clock (scp_server_now) */
} scp_job_info;

typedef struct {
    int64_t queued, active, done, failed;   /* done: finished, result not taken yet */
    int64_t bytes_in_use;                   /* device bytes of admitted and finished jobs */
    int64_t site_updates;                   /* site-steps advanced so far */
    double busy_seconds;                    /* time spent in batched chains */
    double now;
} scp_server_stats;

/* hbm_budget: device bytes jobs may hold at once; quantum: steps per round;
 * max_active: 0 = no limit besides memory. */
int  scp_server_create(int64_t hbm_budget, int quantum, int max_active, scp_server **out);
int  scp_server_destroy(scp_server *s);
int  scp_server_submit(scp_server *s, const scp_job *job, int64_t *id);
/* One scheduling round (admit, step, retire); *finished = jobs retired. */
int  scp_server_round(scp_server *s, int *finished);
int  scp_server_job(const scp_server *s, int64_t id, scp_job_info *info);
/* A finished job's state (cols may be NULL) and diagnostics row (diag may be
 * NULL); frees its device memory. A failed job reports its failure here. */
int  scp_server_take(scp_server *s, int64_t id, double *const *cols, int ncols, double *diag, int ndiag);
int  scp_server_stats_get(const scp_server *s, scp_server_stats *st);
double scp_server_now(const scp_server *s);
/* Rounds a queued job that does not fit may be passed over by later jobs (>= 0). */
int  scp_server_set_aging(scp_server *s, int rounds);

/* Fixed-charge gradient flow: the relaxer sfa/seed/eta_qflow.c (its fixed-Q
 * mode, eta_qflow.c:242-375, and its fixed-(Q, J) mode, below) on the device. A complex-mode state of 12 fields
 * (u_a, v_a = Re, Im Phi_a; tu_a, tv_a = Re, Im Theta_a; a = 0..2) flows by
 * x <- x + dtau F with F the fixed-charge forces (7-point Laplacian, central
 * curl, periodic), omega = Q / N recomputed from the current state every
 * iteration (N = int sum_a |Phi_a|^2 + |Theta_a|^2). The seed (profile, Theta
 * partner, winding) is built on the host (libscp_io: qflow_seed.h). */
typedef struct scp_qflow scp_qflow;

typedef struct {
    int32_t N;            /* sites per edge (periodic cube) */
    int32_t reserved;     /* 0 */
    double L;             /* dx = 2L/(N-1) */
    double eta, m, mth, mu, kappa;
} scp_qflow_params;

/* One printed row of the relaxer (eta_qflow.c:370): iteration, multiplier
 * omega = Q/N of that iteration, E_Q of the updated state, max |F| over the
 * u and tu forces, s_max of the updated state, N of that iteration. */
typedef struct {
    int32_t iter;
    int32_t reserved;
    double omega, E_Q, maxF, s_max, N;
} scp_qflow_row;

int  scp_qflow_create(const scp_qflow_params *p, scp_qflow **out);
int  scp_qflow_destroy(scp_qflow *f);
/* ncols = 12: u0..2, v0..2, tu0..2, tv0..2 (site i*N*N + j*N + k) */
int  scp_qflow_upload(scp_qflow *f, const double *const *cols, int ncols);
int  scp_qflow_download(scp_qflow *f, double *const *cols, int ncols);
/* int sum_a |Phi_a|^2 and int sum_a |Theta_a|^2 of the current state (times dV). */
int  scp_qflow_norm(scp_qflow *f, double *n_phi, double *n_theta);
/* Iterations it = 0..iters at charge Q, step dtau; rows for it % printev == 0
 * and it == iters (the first maxrows kept, *nrows = all); *omega_last: the
 * multiplier of the last iteration (the rotating seed's velocities use it). */
int  scp_qflow_run(scp_qflow *f, double Q, double dtau, int iters, int printev, scp_qflow_row *rows, int maxrows,
                   int *nrows, double *omega_last);

/* One printed row of the fixed-(Q, J) mode (eta_qflow.c:353-357): omega and
 * Omega of that iteration, E_QJ of the updated state, max |F|, s_max, N of that
 * iteration, J_act = omega ell - Omega K, cond = (N K - ell^2) / (N K). */
typedef struct {
    int32_t iter;
    int32_t reserved;
    double omega, Omega, E_QJ, maxF, s_max, N, J_act, cond;
} scp_qflow_row_j;

/* The fixed-(Q, J) flow (eta_qflow.c fixJ = 1, :207-262, :294-310, :377-399):
 * every iteration forms Jhat w = x d_y w - y d_x w - zhat x w of the 12 fields,
 * N, ell = int (u.Jv - v.Ju + tu.Jtv - tv.Jtu), K = int |Jw|^2, the multipliers
 * (omega, Omega) of Q = omega N - Omega ell, J = omega ell - Omega K (Tikhonov
 * 1e-10, |Omega| <= 0.15), and adds the Coriolis and centrifugal forces masked
 * at the light cylinder. Rows as scp_qflow_run. last[0..5) (may be NULL):
 * omega, Omega, cond of the last iteration, ell and K of the final state. */
int  scp_qflow_run_fixj(scp_qflow *f, double Q, double J, double dtau, int iters, int printev,
                        scp_qflow_row_j *rows, int maxrows, int *nrows, double *last);
/* Jhat of the current state's 12 fields (ncols = 12), after scp_qflow_run_fixj
 * (the rotating seed's velocities carry -Omega wm Jhat, eta_qflow.c:447-460);
 * an error if the state changed since. */
int  scp_qflow_download_jhat(scp_qflow *f, double *const *cols, int ncols);

#ifdef __cplusplus
}
#endif
#endif
