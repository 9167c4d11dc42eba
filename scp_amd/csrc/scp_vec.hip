// scp_vec.hip — tricubic patch fits of the real-mode vector frames.
//
// The reference fits each bs^3 patch of each of the six fields (phi_0..2,
// theta_0..2, first rounded to FP32) with the 64 Chebyshev products
// T_a(x) T_b(y) T_c(z), by least squares on the full 64 x 64 normal equations
// (fit_patch, scp_sim.c:2567-2640). The patch samples lie on a tensor grid,
// so the normal matrix is the Kronecker cube of the 4 x 4 one-dimensional
// matrix V^T V and the solution is the separable projection
//     c[a][b][c] = sum_{i,j,k} P[a][i] P[b][j] P[c][k] v[i][j][k],
// P = (V^T V)^{-1} V^T (4 x bs, formed on the host in FP64), the same numbers
// up to FP64 rounding, which the final cast to FP32 hides. One block per
// patch; the three contractions (k, then j, then i) run in FP64 from LDS.
#include "scp_kernels.h"

#include <hip/hip_runtime.h>

namespace scp {

constexpr int VEC_BS_MAX = 16;
constexpr int VEC_THREADS = 256;

struct VecProj {
    double P[4][VEC_BS_MAX];   // P[a][d]
};

__global__ void __launch_bounds__(VEC_THREADS)
k_vec_fit(const double *xo, uint32_t fs, int N, int bs, int bn, VecProj pr, float *out) {
    __shared__ double V[VEC_BS_MAX * VEC_BS_MAX * VEC_BS_MAX];
    __shared__ double T1[VEC_BS_MAX * VEC_BS_MAX * 4];   // [di][dj][c]
    __shared__ double T2[VEC_BS_MAX * 16];               // [di][b][c]
    const int p = blockIdx.x;
    const int i0 = (p / (bn * bn)) * bs, j0 = ((p / bn) % bn) * bs, k0 = (p % bn) * bs;
    const int b3 = bs * bs * bs, b2 = bs * bs;
    const size_t NN = (size_t)N * N;
    for (int f = 0; f < 6; f++) {                        // real mode: u0..2 (phi), tu0..2 (theta)
        const double *src = xo + (size_t)f * fs;
        for (int t = threadIdx.x; t < b3; t += VEC_THREADS) {
            const int di = t / b2, dj = (t / bs) % bs, dk = t % bs;
            V[t] = (double)(float)src[(size_t)(i0 + di) * NN + (size_t)(j0 + dj) * N + (k0 + dk)];
        }
        __syncthreads();
        for (int t = threadIdx.x; t < b2 * 4; t += VEC_THREADS) {   // along k
            const int c = t & 3, ij = t >> 2;
            double s = 0.0;
            for (int dk = 0; dk < bs; dk++) s += pr.P[c][dk] * V[ij * bs + dk];
            T1[t] = s;
        }
        __syncthreads();
        for (int t = threadIdx.x; t < bs * 16; t += VEC_THREADS) {  // along j
            const int c = t & 3, b = (t >> 2) & 3, di = t >> 4;
            double s = 0.0;
            for (int dj = 0; dj < bs; dj++) s += pr.P[b][dj] * T1[(di * bs + dj) * 4 + c];
            T2[t] = s;
        }
        __syncthreads();
        if (threadIdx.x < 64) {                                       // along i
            const int a = threadIdx.x >> 4, bc = threadIdx.x & 15;
            double s = 0.0;
            for (int di = 0; di < bs; di++) s += pr.P[a][di] * T2[di * 16 + bc];
            out[(size_t)p * 384 + f * 64 + threadIdx.x] = (float)s;     // index a*16 + b*4 + c
        }
        __syncthreads();
    }
}

// P = (V^T V)^{-1} V^T for the Chebyshev basis T0..T3 at s = 2 d/(bs-1) - 1,
// d = 0..bs-1; the 4 x 4 inverse by Gauss-Jordan with partial pivoting.
static void vec_projection(int bs, VecProj &pr) {
    double V[VEC_BS_MAX][4], A[4][8];
    for (int d = 0; d < bs; d++) {
        const double s = 2.0 * ((double)d / (bs - 1)) - 1.0;
        V[d][0] = 1.0; V[d][1] = s; V[d][2] = 2.0 * s * s - 1.0; V[d][3] = 4.0 * s * s * s - 3.0 * s;
    }
    for (int a = 0; a < 4; a++)
        for (int b = 0; b < 4; b++) {
            double s = 0.0;
            for (int d = 0; d < bs; d++) s += V[d][a] * V[d][b];
            A[a][b] = s;
            A[a][4 + b] = a == b ? 1.0 : 0.0;
        }
    for (int c = 0; c < 4; c++) {
        int piv = c;
        for (int r = c + 1; r < 4; r++)
            if (__builtin_fabs(A[r][c]) > __builtin_fabs(A[piv][c])) piv = r;
        for (int q = 0; q < 8; q++) { const double t = A[c][q]; A[c][q] = A[piv][q]; A[piv][q] = t; }
        const double inv = 1.0 / A[c][c];
        for (int q = 0; q < 8; q++) A[c][q] *= inv;
        for (int r = 0; r < 4; r++) {
            if (r == c) continue;
            const double m = A[r][c];
            for (int q = 0; q < 8; q++) A[r][q] -= m * A[c][q];
        }
    }
    for (int a = 0; a < 4; a++)
        for (int d = 0; d < VEC_BS_MAX; d++) {
            double s = 0.0;
            if (d < bs)
                for (int b = 0; b < 4; b++) s += A[a][4 + b] * V[d][b];
            pr.P[a][d] = s;
        }
}

int launch_vec_fit(const double *xo, uint32_t fs, int N, int bs, float *out, hipStream_t s) {
    if (bs < 4 || bs > VEC_BS_MAX || N < bs) return (int)hipErrorInvalidValue;
    const int bn = N / bs;
    VecProj pr;
    vec_projection(bs, pr);
    hipLaunchKernelGGL(k_vec_fit, dim3(bn * bn * bn), dim3(VEC_THREADS), 0, s, xo, fs, N, bs, bn, pr, out);
    return (int)hipGetLastError();
}

}  // namespace scp
