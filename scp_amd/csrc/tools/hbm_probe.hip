// hbm_probe.hip — copy ceilings of the HBM for the access patterns of the
// gauged FP32C step (a measurement aid; not part of the library).
//
// Prints, in GB/s and as a fraction of the 8 TB/s peak:
//   * a plain copy (16 bytes per lane, nontemporal), read-only and write-only
//     streams: what the memory delivers to a well-formed stream;
//   * 34 + 34 streams of 512^3 FP32 words read and written site-linearly, one
//     word per lane and stream (the step's 272 B/site: 15 position and 15
//     velocity fields and 4 packed compensation-word fields each way), with the
//     fields at a power-of-two stride (2^27 words) and staggered by 4 KiB;
//   * the same streams in the step's march pattern: 64 x 4 site tiles marching
//     along i over DI planes (the step's XCD-aware tile numbering), at the
//     step's two blocks per CU and with its one-plane-ahead loads (the step's
//     halo re-reads come on top of these streams), in the step's field-major
//     layout (also with the fields staggered by 4 KiB) and with the rows of
//     each array's fields side by side.
// Usage: hbm_probe [reps]   (about 40 GB of device memory, 10-20 s)
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#define HCHK(x)                                                                                  \
    do {                                                                                         \
        hipError_t e_ = (x);                                                                     \
        if (e_ != hipSuccess) {                                                                  \
            fprintf(stderr, "hbm_probe: %s at line %d\n", hipGetErrorString(e_), __LINE__);      \
            exit(1);                                                                             \
        }                                                                                        \
    } while (0)

typedef float f4 __attribute__((ext_vector_type(4)));

__global__ void k_copy4(const f4 *__restrict__ s, f4 *__restrict__ d, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) __builtin_nontemporal_store(__builtin_nontemporal_load(&s[i]), &d[i]);
}
__global__ void k_read4(const f4 *__restrict__ s, float *__restrict__ sink, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const f4 v = __builtin_nontemporal_load(&s[i]);
        if (v.x + v.y + v.z + v.w == 1234.5f) sink[0] = v.x;   // never true: keeps the load
    }
}
__global__ void k_write4(f4 *__restrict__ d, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) __builtin_nontemporal_store(f4{1.f, 2.f, 3.f, 4.f}, &d[i]);
}

constexpr int NFP = 15, NQ = 4, NS = 2 * NFP + NQ;   // streams each way
struct Streams {
    const float *in[NS];
    float *out[NS];
};

__global__ void k_linear(Streams p, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float v[NS];
#pragma unroll
    for (int f = 0; f < NS; f++) v[f] = __builtin_nontemporal_load(p.in[f] + i);
#pragma unroll
    for (int f = 0; f < NS; f++) __builtin_nontemporal_store(v[f] * 1.0001f, p.out[f] + i);
}

__device__ inline uint32_t xcd_block(uint32_t b, uint32_t nb) { return (b % 8) * (nb / 8) + b / 8; }

// march pattern over an N^3 grid (N a multiple of 64): each block streams the
// own sites of one 64 x 4 tile along i; like the step, the loads of plane i+1
// are issued before the stores of plane i. No halo: the step's halo re-reads
// come on top of these streams.
// Strides within a stream (each stream has its own base pointer): row and
// plane strides of the position and velocity streams (rs, ps) and of the
// word streams (rsq, psq). Field-major: N and N^2 for all. Row-interleaved
// (the rows of one array's fields side by side): 15 N and 15 N^2, 4 N and 4 N^2.
struct Layout {
    size_t rs, ps, rsq, psq;
};
__global__ void __launch_bounds__(256) k_march_pattern(Streams p, int N, int DI, Layout L) {
    const uint32_t ntk = N / 64, ntj = N / 4;
    uint32_t lb = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t tk = lb % ntk;
    lb /= ntk;
    const uint32_t tj = lb % ntj, seg = lb / ntj;
    const int k = tk * 64 + threadIdx.x, j = tj * 4 + threadIdx.y;
    const int i0 = seg * DI, i1 = min(N, i0 + DI);
    extern __shared__ float lds_reserved[];   // only caps the blocks per CU
    if (N < 0) lds_reserved[threadIdx.x] = 0.f;
    auto load = [&](int i, float *v) {
        const size_t e = (size_t)i * L.ps + (size_t)j * L.rs + k, eq = (size_t)i * L.psq + (size_t)j * L.rsq + k;
#pragma unroll
        for (int f = 0; f < NS; f++)
            v[f] = f < NFP ? p.in[f][e] : __builtin_nontemporal_load(p.in[f] + (f < 2 * NFP ? e : eq));
    };
    float v[NS];
    if (i0 < i1) load(i0, v);
    for (int i = i0; i < i1; i++) {
        float vn[NS] = {};
        if (i + 1 < i1) load(i + 1, vn);
        const size_t e = (size_t)i * L.ps + (size_t)j * L.rs + k, eq = (size_t)i * L.psq + (size_t)j * L.rsq + k;
#pragma unroll
        for (int f = 0; f < NS; f++) __builtin_nontemporal_store(v[f] * 1.0001f, p.out[f] + (f < 2 * NFP ? e : eq));
#pragma unroll
        for (int f = 0; f < NS; f++) v[f] = vn[f];
    }
}

// the FP32C step's staged planes per block: 3 planes x 15 fields x 6 x 66 words
constexpr size_t LDS_STEP = 3 * 15 * 6 * 66 * 4;

static hipEvent_t ev0, ev1;
template <class F>
static double time_ms(F launch, int reps) {
    launch();
    HCHK(hipGetLastError());
    HCHK(hipDeviceSynchronize());
    HCHK(hipEventRecord(ev0));
    for (int r = 0; r < reps; r++) launch();
    HCHK(hipEventRecord(ev1));
    HCHK(hipEventSynchronize(ev1));
    float ms = 0.f;
    HCHK(hipEventElapsedTime(&ms, ev0, ev1));
    return ms / reps;
}
static void line(const char *what, double bytes, double ms) {
    const double gbs = bytes / ms / 1e6;
    printf("%-62s %8.3f ms %8.1f GB/s %6.1f %% of 8 TB/s\n", what, ms, gbs, gbs / 80.0);
    fflush(stdout);
}

int main(int argc, char **argv) {
    const int reps = argc > 1 ? atoi(argv[1]) : 10;
    HCHK(hipEventCreate(&ev0));
    HCHK(hipEventCreate(&ev1));
    {
        const size_t n4 = (size_t)1 << 28;   // 4 GiB of 16-byte vectors each way
        f4 *s, *d;
        float *sink;
        HCHK(hipMalloc(&s, n4 * 16));
        HCHK(hipMalloc(&d, n4 * 16));
        HCHK(hipMalloc(&sink, 16));
        HCHK(hipMemset(s, 0, n4 * 16));
        const uint32_t nb = (uint32_t)(n4 / 256);
        line("copy, 16 B per lane, nontemporal (4 GiB each way)", 2.0 * n4 * 16,
             time_ms([&] { k_copy4<<<nb, 256>>>(s, d, n4); }, reps));
        line("read only, 16 B per lane (4 GiB)", 1.0 * n4 * 16, time_ms([&] { k_read4<<<nb, 256>>>(s, sink, n4); }, reps));
        line("write only, 16 B per lane (4 GiB)", 1.0 * n4 * 16, time_ms([&] { k_write4<<<nb, 256>>>(d, n4); }, reps));
        HCHK(hipFree(s));
        HCHK(hipFree(d));
        HCHK(hipFree(sink));
    }
    const int N = 512;
    const size_t n3 = (size_t)N * N * N;
    const double bytes = 2.0 * NS * 4 * n3;   // 272 B per site
    for (int stagger = 0; stagger < 2; stagger++) {
        const size_t fstride = n3 + (stagger ? 1024 : 0);   // words between fields
        float *slab;
        HCHK(hipMalloc(&slab, 2 * NS * fstride * 4));
        HCHK(hipMemset(slab, 0, 2 * NS * fstride * 4));
        Streams p;
        for (int f = 0; f < NS; f++) {
            p.in[f] = slab + f * fstride;
            p.out[f] = slab + (NS + f) * fstride;
        }
        char what[96];
        snprintf(what, sizeof what, "34 + 34 streams, site-linear, fields %s", stagger ? "staggered by 4 KiB" : "2^27 words apart");
        line(what, bytes, time_ms([&] { k_linear<<<(uint32_t)(n3 / 256), 256>>>(p, n3); }, reps));
        {
            for (int DI : {128, 32}) {
                if (stagger && DI != 32) continue;
                const uint32_t nb = (uint32_t)((N / 64) * (N / 4) * ((N + DI - 1) / DI));
                // LDS reserved as the step's staged planes (71 KB per block): two blocks per CU, as in the step
                const Layout fm{(size_t)N, (size_t)N * N, (size_t)N, (size_t)N * N};
                snprintf(what, sizeof what, "34 + 34 streams, march pattern 64 x 4 x %d planes%s", DI,
                         stagger ? ", staggered" : "");
                line(what, bytes, time_ms([&] { k_march_pattern<<<nb, dim3(64, 4), LDS_STEP>>>(p, N, DI, fm); }, reps));
            }
        }
        HCHK(hipFree(slab));
    }
    {
        // row-interleaved step state: in each array (positions, velocities, words;
        // in and out apart) the rows of its fields for one (i, j) side by side
        float *slab;
        HCHK(hipMalloc(&slab, 2 * NS * n3 * 4));
        HCHK(hipMemset(slab, 0, 2 * NS * n3 * 4));
        Streams p;
        const int first[3] = {0, NFP, 2 * NFP}, count[3] = {NFP, NFP, NQ};
        size_t base = 0;
        for (int side = 0; side < 2; side++)
            for (int a = 0; a < 3; a++) {
                for (int f = 0; f < count[a]; f++) {
                    float *q = slab + base + (size_t)f * N;
                    if (side == 0) p.in[first[a] + f] = q;
                    else p.out[first[a] + f] = q;
                }
                base += (size_t)count[a] * n3;
            }
        const Layout ri{(size_t)NFP * N, (size_t)NFP * N * N, (size_t)NQ * N, (size_t)NQ * N * N};
        const int DI = 32;
        const uint32_t nb = (uint32_t)((N / 64) * (N / 4) * ((N + DI - 1) / DI));
        line("34 + 34 streams, march pattern 64 x 4 x 32 planes, row-interleaved", bytes,
             time_ms([&] { k_march_pattern<<<nb, dim3(64, 4), LDS_STEP>>>(p, N, DI, ri); }, reps));
        HCHK(hipFree(slab));
    }
    return 0;
}
