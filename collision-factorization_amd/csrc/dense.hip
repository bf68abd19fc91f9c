// Dense-tensor contraction (dense_tensor_kernel_mex) and the dense tensor's assembly on the device.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "af_device.hpp"
#include "common.hpp"

namespace cfk {

// ---------------------------------------------------------------------------------------------
// Dense (dense_tensor_kernel_mex.cpp:14-29): Q[i] += sum_{j,k} C[i + jN + kN^2] f[j] f[k].
// C is N^2 rows r = j + kN of N contiguous doubles, weighted by w[r] = f[j] f[k]: a stream of N^3
// doubles read once per call (HBM-bound).  One workgroup streams a contiguous range of rows (one round
// of the GPU's workgroup slots covers C); each thread owns V consecutive columns (V = 2: one 16-byte load per row when N is even) and keeps their
// sums for the NB distributions of the call in registers.  UNR rows are requested before the first is
// used, so UNR loads per lane are in flight.  f of the call's distributions sits in LDS (read as a
// broadcast, the row index being uniform).  Each workgroup writes its partial sums, which two small
// kernels add in workgroup order (deterministic).
// ---------------------------------------------------------------------------------------------
typedef double dense_f64x2 __attribute__((ext_vector_type(2)));

template <int V, int NB, int UNR>
__global__ __launch_bounds__(1024) void dense_rows_kernel(const double *__restrict__ C, const double *__restrict__ f,
                                                          double *__restrict__ part, int64_t N, int nb,
                                                          int64_t rows_per_blk) {
    extern __shared__ double fs[];  // [nb][N]
    for (int64_t e = threadIdx.x; e < (int64_t)nb * N; e += blockDim.x) fs[e] = f[e];
    __syncthreads();
    const int64_t nvec = N / V;
    const int64_t cv = (int64_t)blockIdx.y * blockDim.x + threadIdx.x;  // this thread's column vector
    const bool on = cv < nvec;
    const int64_t r0 = (int64_t)blockIdx.x * rows_per_blk;
    const int64_t r1 = min(N * N, r0 + rows_per_blk);
    double acc[NB][V];
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
        for (int v = 0; v < V; ++v) acc[b][v] = 0.0;
    const double *col = C + (on ? cv * V : 0);
    int64_t j = r0 % N, k = r0 / N;  // row r = j + kN, advanced with r (wave-uniform)
    for (int64_t r = r0; r < r1; r += UNR) {
        double c[UNR][V];
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
            if (r + u < r1 && on) {
                if constexpr (V == 2) {
                    const dense_f64x2 x = __builtin_nontemporal_load(reinterpret_cast<const dense_f64x2 *>(col + (r + u) * N));
                    c[u][0] = x[0];
                    c[u][1] = x[1];
                } else {
                    c[u][0] = __builtin_nontemporal_load(col + (r + u) * N);
                }
            } else {
#pragma unroll
                for (int v = 0; v < V; ++v) c[u][v] = 0.0;
            }
        }
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
            if (r + u < r1) {
#pragma unroll
                for (int b = 0; b < NB; ++b) {
                    if (b < nb) {
                        const double w = fs[b * N + j] * fs[b * N + k];
#pragma unroll
                        for (int v = 0; v < V; ++v) acc[b][v] = fma(c[u][v], w, acc[b][v]);
                    }
                }
                if (++j == N) {
                    j = 0;
                    ++k;
                }
            }
        }
    }
    if (on) {
#pragma unroll
        for (int b = 0; b < NB; ++b)
            if (b < nb)
#pragma unroll
                for (int v = 0; v < V; ++v) part[((int64_t)blockIdx.x * nb + b) * N + cv * V + v] = acc[b][v];
    }
}

// Dense tensor from the COO and R (GeneralCollisionTensor.m:283-311, assemble_full_tensor):
// C[(k1 N_Q + q1) + (k2 N_Q + q2) N + (k3 N_Q + q3) N^2] = R_t[k1,k2,k3] g_z, one workgroup per nonzero z.
// Distinct nonzeros address distinct entries, so the assignment order does not matter; every other
// entry is zero (C is cleared first).
__global__ __launch_bounds__(256) void dense_assemble_kernel(double *__restrict__ C, const int4 *__restrict__ zl,
                                                              const double *__restrict__ zg, const double *__restrict__ R,
                                                              int N_Q, int K) {
    const int4 e = zl[blockIdx.x];  // (q1, q2, q3, t), 0-based
    const double g = zg[blockIdx.x];
    const int64_t N = (int64_t)N_Q * K, K3 = (int64_t)K * K * K;
    const double *Rt = R + (int64_t)e.w * K3;
    for (int64_t x = threadIdx.x; x < K3; x += blockDim.x) {
        const int k1 = (int)(x % K), k2 = (int)((x / K) % K), k3 = (int)(x / ((int64_t)K * K));
        C[((int64_t)k1 * N_Q + e.x) + ((int64_t)k2 * N_Q + e.y) * N + ((int64_t)k3 * N_Q + e.z) * N * N] = Rt[x] * g;
    }
}

__global__ void fill_zero_kernel(double *__restrict__ p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = 0.0;
}

// Sum of the workgroups' partial sums, in two levels so that no thread walks all of them: level 1 adds
// runs of DENSE_RUN consecutive workgroups' partials (one thread per output and run), level 2 adds the
// runs in order and accumulates into Q.  Fixed summation order: deterministic.
constexpr int DENSE_RUN = 32;

__global__ __launch_bounds__(256) void dense_reduce_runs_kernel(const double *__restrict__ part, double *__restrict__ part2,
                                                                 int64_t n, int nchunk) {
    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= n) return;
    const int c0 = blockIdx.y * DENSE_RUN, c1 = min(nchunk, c0 + DENSE_RUN);
    double s = 0.0;
    for (int ch = c0; ch < c1; ++ch) s += part[(int64_t)ch * n + idx];
    part2[(int64_t)blockIdx.y * n + idx] = s;
}

__global__ __launch_bounds__(256) void dense_reduce_kernel(const double *__restrict__ part2, double *__restrict__ Q,
                                                            int64_t n, int nrun) {
    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= n) return;
    double s = 0.0;
    for (int r = 0; r < nrun; ++r) s += part2[(int64_t)r * n + idx];
    Q[idx] += s;
}

// Launch shape of the dense stream for N: distributions per pass, threads per workgroup, column tiles
// and row chunks (workgroups along the rows: one round of the GPU's workgroup slots, at least 16 rows each).
struct DenseShape {
    int nbmax, threads, ncol, nchunk, vec;
};

static int dense_shape(int64_t N, DenseShape &sh) {
    sh.vec = (N % 2 == 0) ? 2 : 1;
    sh.nbmax = N <= 1024 ? 4 : 1;  // f of the pass in LDS: nb * N doubles <= 32 KiB
    const int64_t nvec = N / sh.vec;
    sh.threads = (int)std::min<int64_t>(1024, (nvec + WAVE - 1) / WAVE * WAVE);
    sh.ncol = (int)((nvec + sh.threads - 1) / sh.threads);
    int dev = 0, ncu = 0, per_cu = 0;
    CFK_HIP(hipGetDevice(&dev));
    CFK_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const size_t lds = (size_t)sh.nbmax * N * sizeof(double);
    if (sh.vec == 2) {
        CFK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, dense_rows_kernel<2, 4, 8>, sh.threads, lds));
    } else {
        CFK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, dense_rows_kernel<1, 4, 8>, sh.threads, lds));
    }
    const int64_t slots = (int64_t)std::max(1, per_cu) * std::max(1, ncu);
    const int64_t by_rows = std::max<int64_t>(1, N * N / 16);
    sh.nchunk = (int)std::max<int64_t>(1, std::min<int64_t>(by_rows, slots / sh.ncol));
    return CFK_OK;
}

// Workspace of one call in doubles: the workgroups' partial sums, then the level-1 run sums.
int dense_part_doubles(int64_t N, int64_t *n) {
    DenseShape sh;
    int rc;
    if ((rc = dense_shape(N, sh))) return rc;
    const int64_t nrun = (sh.nchunk + DENSE_RUN - 1) / DENSE_RUN;
    *n = ((int64_t)sh.nchunk + nrun) * sh.nbmax * N;
    return CFK_OK;
}

int launch_dense(double *Q, const double *f, const double *C, int64_t N, int64_t B, double *part, hipStream_t st) {
    DenseShape sh;
    int rc;
    if ((rc = dense_shape(N, sh))) return rc;
    const int64_t rows_per_blk = (N * N + sh.nchunk - 1) / sh.nchunk;
    for (int64_t b0 = 0; b0 < B; b0 += sh.nbmax) {
        const int nb = (int)std::min<int64_t>(sh.nbmax, B - b0);
        const dim3 grid((unsigned)sh.nchunk, (unsigned)sh.ncol);
        const size_t lds = (size_t)nb * N * sizeof(double);
        if (sh.vec == 2)
            hipLaunchKernelGGL((dense_rows_kernel<2, 4, 8>), grid, dim3(sh.threads), lds, st, C, f + b0 * N, part, N, nb,
                               rows_per_blk);
        else
            hipLaunchKernelGGL((dense_rows_kernel<1, 4, 8>), grid, dim3(sh.threads), lds, st, C, f + b0 * N, part, N, nb,
                               rows_per_blk);
        CFK_HIP(hipGetLastError());
        const int64_t tot = N * nb;
        const int nrun = (sh.nchunk + DENSE_RUN - 1) / DENSE_RUN;
        double *part2 = part + (int64_t)sh.nchunk * sh.nbmax * N;
        hipLaunchKernelGGL(dense_reduce_runs_kernel, dim3((unsigned)((tot + 255) / 256), (unsigned)nrun), dim3(256), 0, st,
                           part, part2, tot, sh.nchunk);
        CFK_HIP(hipGetLastError());
        hipLaunchKernelGGL(dense_reduce_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, part2, Q + b0 * N,
                           tot, nrun);
        CFK_HIP(hipGetLastError());
    }
    return CFK_OK;
}

int launch_dense_assemble(double *C, const int4 *zl, const double *zg, const double *R, int64_t N_G, int N_Q, int K,
                          hipStream_t st) {
    const int64_t N = (int64_t)N_Q * K, n = N * N * N;
    hipLaunchKernelGGL(fill_zero_kernel, dim3((unsigned)std::min<int64_t>((n + 255) / 256, 65536)), dim3(256), 0, st, C, n);
    CFK_HIP(hipGetLastError());
    if (N_G > 0) {
        hipLaunchKernelGGL(dense_assemble_kernel, dim3((unsigned)N_G), dim3(256), 0, st, C, zl, zg, R, N_Q, K);
        CFK_HIP(hipGetLastError());
    }
    return CFK_OK;
}

}  // namespace cfk
