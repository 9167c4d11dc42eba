// scp_grid.h — internal: the device state behind the opaque scp_grid handle
// and the helpers the C-ABI files share (scp_abi.hip, scp_batch.hip).
#pragma once

#include "scp_amd.h"
#include "scp_kernels.h"

#include <hip/hip_runtime.h>

namespace scp_internal {
// device memory source of a grid (nullptr: hipMalloc / hipFree); the server
// serves its grids from a pool through this
struct GridMem {
    void *ctx;
    hipError_t (*alloc)(void *ctx, void **p, size_t bytes);
    void (*release)(void *ctx, void *p);
};
}  // namespace scp_internal

struct scp_grid {
    scp_params p;
    const scp_internal::GridMem *mem;   // allocator of every device buffer below (null: hipMalloc)
    scp::KParams kp;
    int mode;        // storage mode (0/1/2), number of columns follows it
    int step_mode;   // kernel mode: 2 only when gauged with g != 0
    int NP;
    hipStream_t stream;
    int device;
    int f32;          // SCP_PREC_F32C
    int lo_kind;      // LoKind of the compensation words
    double *P[2];     // FP64 state
    double *W;
    float *P32[2];    // FP32C state: hi words ...
    float *W32;
    void *Plo[2];     // ... and compensation words
    void *Wlo;
    double *xo, *vo;
    double *aux;     // real-mode pass-1 fields (Cosserat, hardening, theta frequency: 7), when needed
    double *acc;     // forces output (lazily allocated)
    double *partials;
    double *dout;
    double *rho;      // fused diagnostics: Gauss charge density (N^3, gauged), lazily allocated
    double *fpart;    // fused diagnostics: partial rows of the step and the Gauss tail, lazily allocated
    int fpart_rows;
    double *pinx, *pinv;   // bc 1: pinned positions / velocities (emitted layout)
    double *side;          // bc 1: side rows of the y-edge pass
    unsigned long long *argmin;
    int cur;
    int nblocks_diag;
    int64_t steps;
    hipEvent_t ev0, ev1;
    char kname[96];
    struct scp_batch *batch;   // the batch this grid belongs to, if any
};

namespace scp {
// device seeding (scp_seed.hip): fills xo/vo (device field order) of an N^3
// grid with field stride fs; synchronous on s; returns a hipError_t
int seed_balls_device(int N, double L, double dx, uint32_t fs, int gauged, const scp_qball *balls, int nballs,
                      double theta_sigma, double link_sigma, uint64_t seed, double *xo, double *vo,
                      hipStream_t s, const scp_internal::GridMem *mem);
}  // namespace scp

namespace scp_internal {
int set_err(int code, const char *fmt, ...);
int kerr(int rc, const char *what);
// K_STEP runs the LDS-staged march for this grid
bool use_march(const scp_grid *g, int kind);
scp::Bufs<double> bufs_for(scp_grid *g, int kind);
scp::Bufs32 bufs32_for(scp_grid *g, int kind);
// initial force + first half kick + drift from xo/vo (upload and device seeding)
int run_init(scp_grid *g, bool pin = true);
// device buffers of a grid through its allocator
hipError_t gmalloc(scp_grid *g, void **p, size_t bytes);
void gfree(scp_grid *g, void *p);
// scp_grid_create with an allocator; emit = false leaves the FP64 emit buffers
// xo, vo unallocated (grid_emit_buffers adds them when a launch must emit)
int grid_create_with(const scp_params *p, const GridMem *mem, bool emit, scp_grid **out);
// allocate (on) or release (off) xo, vo; steps of a grid without them do not emit
int grid_emit_buffers(scp_grid *g, bool on);
// release the step state (P, W and their words) of a grid whose last step has
// emitted: only xo/vo and the diagnostics buffers remain (no further steps)
int grid_release_step_state(scp_grid *g);
// device bytes of the emit buffers xo, vo and of everything else
int64_t grid_emit_bytes(const scp_params *p);
}  // namespace scp_internal

// return the HIP error of x as status -2 with a message
#define HIPCHK(x)                                                                              \
    do {                                                                                       \
        hipError_t e_ = (x);                                                                   \
        if (e_ != hipSuccess)                                                                  \
            return scp_internal::set_err(-2, "%s failed: %s (%s:%d)", #x, hipGetErrorString(e_), \
                                         __FILE__, __LINE__);                                  \
    } while (0)
