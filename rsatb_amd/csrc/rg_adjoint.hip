// Kernels of the adjoint (reverse-mode) Chebyshev recurrence.
//
// The backward sweep itself reuses the fused SpMM (epilogue EPI_ADJ); what
// is specific to it is the contraction of adjoint and forward states onto
// the operator's sparsity pattern, dL/dA_ij += coef sum_c Re(u_i,c conj v_j,c).
#include "rg_internal.h"

namespace rg {
namespace {

__global__ __launch_bounds__(256) void k_grad_accum(int64_t nbr, int b, int64_t nnzb,
                                                    const int64_t *__restrict__ row_ptr,
                                                    const int32_t *__restrict__ col_ind,
                                                    const double *__restrict__ U,
                                                    const double *__restrict__ V, int R, int cplx,
                                                    double coef, double *__restrict__ G) {
    const int per = b * b;
    const int64_t n = nnzb * per;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t s = t / per;
        const int e = (int)(t - s * per);
        const int cc = e / b, rr = e - cc * b; // column-major element (rr, cc)
        int64_t lo = 0, hi = nbr - 1;          // block row of s
        while (lo < hi) {
            const int64_t mid = (lo + hi + 1) >> 1;
            if (row_ptr[mid] <= s) lo = mid;
            else hi = mid - 1;
        }
        const int64_t i = lo * b + rr;
        const int64_t j = (int64_t)col_ind[s] * b + cc;
        double acc = 0.0;
        if (cplx) {
            const double2 *u = reinterpret_cast<const double2 *>(U) + i * R;
            const double2 *v = reinterpret_cast<const double2 *>(V) + j * R;
            for (int c = 0; c < R; ++c) {
                const double2 a = u[c], w = v[c];
                acc = __builtin_fma(a.x, w.x, __builtin_fma(a.y, w.y, acc));
            }
        } else {
            const double *u = U + i * R;
            const double *v = V + j * R;
            for (int c = 0; c < R; ++c) acc = __builtin_fma(u[c], v[c], acc);
        }
        G[t] += coef * acc;
    }
}

__global__ __launch_bounds__(256) void k_zero(double *__restrict__ y, int64_t n) {
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (int64_t)gridDim.x * blockDim.x)
        y[t] = 0.0;
}

__global__ __launch_bounds__(256) void k_axpby(double *__restrict__ y, const double *__restrict__ x,
                                               double alpha, double beta, int64_t n) {
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (int64_t)gridDim.x * blockDim.x)
        y[t] = alpha * x[t] + beta * y[t];
}

int grid_for(int64_t n) {
    int64_t g = (n + 255) / 256;
    if (g > 16384) g = 16384;
    if (g < 1) g = 1;
    return (int)g;
}

} // namespace

hipError_t launch_grad_accum(int64_t nbr, int b, int64_t nnzb, const int64_t *row_ptr,
                             const int32_t *col_ind, const double *U, const double *V, int R,
                             bool cplx, double coef, double *G, hipStream_t st) {
    hipLaunchKernelGGL(k_grad_accum, dim3(grid_for(nnzb * b * b)), dim3(256), 0, st, nbr, b, nnzb,
                       row_ptr, col_ind, U, V, R, cplx ? 1 : 0, coef, G);
    return hipGetLastError();
}

hipError_t launch_zero(double *y, int64_t n, hipStream_t st) {
    hipLaunchKernelGGL(k_zero, dim3(grid_for(n)), dim3(256), 0, st, y, n);
    return hipGetLastError();
}

hipError_t launch_axpby(double *y, const double *x, double alpha, double beta, int64_t n, hipStream_t st) {
    hipLaunchKernelGGL(k_axpby, dim3(grid_for(n)), dim3(256), 0, st, y, x, alpha, beta, n);
    return hipGetLastError();
}

} // namespace rg
