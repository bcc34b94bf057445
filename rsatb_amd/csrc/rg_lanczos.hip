// Vector kernels of the spectral-bound Lanczos (replaces the PRIMME extreme
// eigenvalue solve of rsatb/setup/spectral_bounds.hpp:311-424): R
// independent Lanczos chains run as the columns of one block, so the
// operator apply is the same fused SpMM chain the moment loop uses and the
// three-term update needs no host round trip.
#include "rg_internal.h"

namespace rg {
namespace {

constexpr int kThreads = 256;

// w <- w - alpha_c v - beta_c vprev;  partials[wg][c] = sum_rows |w|^2.
// R divides 256; element (row, c) lives at row * R + c (complex: 2 doubles).
__global__ __launch_bounds__(kThreads) void k_lanczos_axpy(int64_t rows, int R, int cplx,
                                                           double *__restrict__ w,
                                                           const double *__restrict__ v,
                                                           const double *__restrict__ vprev,
                                                           const double2 *__restrict__ alpha,
                                                           const double2 *__restrict__ bsq,
                                                           double2 *__restrict__ partials) {
    const int c = threadIdx.x % R;
    const int rsub = threadIdx.x / R;
    const int rstep = kThreads / R;
    const double a = alpha[c].x;
    const double bq = bsq[c].x;
    const double bt = bq > 0.0 ? sqrt(bq) : 0.0;
    const int comp = cplx ? 2 : 1;
    double acc = 0.0;
    for (int64_t r = (int64_t)blockIdx.x * rstep + rsub; r < rows; r += (int64_t)gridDim.x * rstep) {
        const int64_t base = (r * R + c) * comp;
        for (int q = 0; q < comp; ++q) {
            const double x = w[base + q] - a * v[base + q] - bt * vprev[base + q];
            w[base + q] = x;
            acc = __builtin_fma(x, x, acc);
        }
    }
    __shared__ double sm[kThreads];
    sm[threadIdx.x] = acc;
    __syncthreads();
    if (threadIdx.x < R) {
        double s = 0.0;
        for (int k = threadIdx.x; k < kThreads; k += R) s += sm[k];
        partials[(int64_t)blockIdx.x * R + threadIdx.x] = double2{s, 0.0};
    }
}

// v <- w / sqrt(bsq_c) (0 where the chain has terminated).
__global__ __launch_bounds__(kThreads) void k_lanczos_scale(int64_t rows, int R, int cplx,
                                                            const double *__restrict__ w,
                                                            const double2 *__restrict__ bsq,
                                                            double tiny, double *__restrict__ v) {
    const int comp = cplx ? 2 : 1;
    const int64_t n = rows * R * comp;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)((t / comp) % R);
        const double bq = bsq[c].x;
        v[t] = bq > tiny ? w[t] / sqrt(bq) : 0.0;
    }
}

} // namespace

int lanczos_grid(int64_t rows, int R) {
    const int64_t rstep = kThreads / R;
    int64_t g = (rows + rstep - 1) / rstep;
    if (g > 1024) g = 1024;
    if (g < 1) g = 1;
    return (int)g;
}

hipError_t launch_lanczos_axpy(int64_t rows, int R, bool cplx, double *w, const double *v,
                               const double *vprev, const double2 *alpha, const double2 *bsq,
                               double2 *partials, int grid, hipStream_t st) {
    hipLaunchKernelGGL(k_lanczos_axpy, dim3(grid), dim3(kThreads), 0, st, rows, R, cplx ? 1 : 0, w,
                       v, vprev, alpha, bsq, partials);
    return hipGetLastError();
}

hipError_t launch_lanczos_scale(int64_t rows, int R, bool cplx, const double *w,
                                const double2 *bsq, double tiny, double *v, hipStream_t st) {
    int64_t n = rows * R * (cplx ? 2 : 1);
    int64_t g = (n + kThreads - 1) / kThreads;
    if (g > 8192) g = 8192;
    hipLaunchKernelGGL(k_lanczos_scale, dim3((unsigned)g), dim3(kThreads), 0, st, rows, R,
                       cplx ? 1 : 0, w, bsq, tiny, v);
    return hipGetLastError();
}

} // namespace rg
