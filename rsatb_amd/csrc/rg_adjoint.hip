// Kernels of the adjoint (reverse-mode) Chebyshev recurrence.
//
// The backward sweep itself reuses the fused SpMM (epilogue EPI_ADJ); what
// is specific to it is the contraction of adjoint and forward states onto
// the operator's sparsity pattern, dL/dA_ij += coef sum_c Re(u_i,c conj v_j,c).
#include "rg_internal.h"

#include <type_traits>

namespace rg {
namespace {

// G[e] += coef * sum_c Re(U[i, c] conj(V[j, c])) for every stored entry e of
// the pattern, and with PAIR2 also + coef * sum_c Re(U2[i, c] conj(V2[j, c]))
// in the same pass. One group of LPR lanes per point row i (lanes over the
// sample columns, c = q + t LPR): the row's U (and U2) values stay in
// registers while the group walks the row's entries, gathering each V (V2)
// row with coalesced loads and summing the lane partials by xor shuffles.
// Entries are owned by their point row, so no two groups write the same G
// element. For b = 1, four entries' rows are loaded before their reductions,
// and their four G values with the column indices (every lane of the group
// loads them; lane 0 writes the sums back), so no load waits on a reduction.
template <bool CPLX, int LPR, int UPL, bool B1, bool PAIR2>
__global__ __launch_bounds__(256) void k_grad_accum(int64_t nbr, int b, const int64_t *__restrict__ row_ptr,
                                                    const int32_t *__restrict__ col_ind,
                                                    const double *__restrict__ U, const double *__restrict__ V,
                                                    const double *__restrict__ U2, const double *__restrict__ V2,
                                                    int R, double coef, double *__restrict__ G) {
    using VT = std::conditional_t<CPLX, double2, double>;
    constexpr int NG = 64 / LPR;
    const VT *__restrict__ Uv = reinterpret_cast<const VT *>(U);
    const VT *__restrict__ Vv = reinterpret_cast<const VT *>(V);
    const VT *__restrict__ Uv2 = reinterpret_cast<const VT *>(U2);
    const VT *__restrict__ Vv2 = reinterpret_cast<const VT *>(V2);
    const int lane = threadIdx.x & 63;
    const int q = lane % LPR;
    const int64_t nrows = nbr * b;
    const int64_t stride = (int64_t)gridDim.x * (blockDim.x >> 6) * NG;
    auto dot = [](VT a, VT w) {
        if constexpr (CPLX) return __builtin_fma(a.x, w.x, a.y * w.y);
        else return a * w;
    };
    auto vzero = []() {
        if constexpr (CPLX) return double2{0.0, 0.0};
        else return 0.0;
    };
    auto reduce = [](double v) {
#pragma unroll
        for (int off = LPR / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, LPR);
        return v;
    };
    for (int64_t i = ((int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * NG + lane / LPR; i < nrows;
         i += stride) {
        const int64_t br = B1 ? i : i / b;
        const int rr = B1 ? 0 : (int)(i - br * b);
        VT u[UPL], u2[UPL];
#pragma unroll
        for (int t = 0; t < UPL; ++t) {
            const int c = q + t * LPR;
            u[t] = c < R ? Uv[i * R + c] : vzero();
            if constexpr (PAIR2) u2[t] = c < R ? Uv2[i * R + c] : vzero();
        }
        const int64_t s0 = row_ptr[br], s1 = row_ptr[br + 1];
        int64_t s = s0;
        if constexpr (B1) {
            constexpr int K = 4;
            for (; s + K <= s1; s += K) {
                double acc[K], g0[K];
                int64_t jk[K];
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    jk[k] = col_ind[s + k];
                    g0[k] = G[s + k];
                }
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    const int64_t j = jk[k];
                    acc[k] = 0.0;
#pragma unroll
                    for (int t = 0; t < UPL; ++t) {
                        const int c = q + t * LPR;
                        if (c < R) acc[k] += dot(u[t], Vv[j * R + c]);
                        if constexpr (PAIR2)
                            if (c < R) acc[k] += dot(u2[t], Vv2[j * R + c]);
                    }
                }
#pragma unroll
                for (int k = 0; k < K; ++k) acc[k] = reduce(acc[k]);
                if (q == 0)
#pragma unroll
                    for (int k = 0; k < K; ++k) G[s + k] = g0[k] + coef * acc[k];
            }
        }
        for (; s < s1; ++s) {
            const int64_t jb = (int64_t)col_ind[s] * b;
            for (int cc = 0; cc < b; ++cc) {
                double acc = 0.0;
#pragma unroll
                for (int t = 0; t < UPL; ++t) {
                    const int c = q + t * LPR;
                    if (c < R) acc += dot(u[t], Vv[(jb + cc) * R + c]);
                    if constexpr (PAIR2)
                        if (c < R) acc += dot(u2[t], Vv2[(jb + cc) * R + c]);
                }
                acc = reduce(acc);
                // column-major element (rr, cc) of block s
                if (q == 0) G[s * b * b + (int64_t)cc * b + rr] += coef * acc;
            }
        }
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

// out[c * rows + r] = in[c * ld_in + r] widened to double, r < rows, c < cols
__global__ __launch_bounds__(256) void k_widen_cols(const float *__restrict__ in, int64_t ld_in,
                                                    double *__restrict__ out, int64_t rows, int cols) {
    const int64_t n = rows * cols;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t c = t / rows, r = t - c * rows;
        out[t] = (double)in[c * ld_in + r];
    }
}

int grid_for(int64_t n) {
    int64_t g = (n + 255) / 256;
    if (g > 16384) g = 16384;
    if (g < 1) g = 1;
    return (int)g;
}

} // namespace

template <bool CPLX, int LPR, int UPL>
hipError_t go_grad(int64_t nbr, int b, const int64_t *row_ptr, const int32_t *col_ind, const double *U,
                   const double *V, const double *U2, const double *V2, int R, double coef, double *G,
                   hipStream_t st) {
    const int64_t rows_per_wg = 4 * (64 / LPR);
    int64_t g = (nbr * b + rows_per_wg - 1) / rows_per_wg;
    if (g > 16384) g = 16384;
    if (g < 1) g = 1;
    const dim3 gd((unsigned)g), bd(256);
    if (U2) {
        if (b == 1)
            hipLaunchKernelGGL((k_grad_accum<CPLX, LPR, UPL, true, true>), gd, bd, 0, st, nbr, b, row_ptr, col_ind, U,
                               V, U2, V2, R, coef, G);
        else
            hipLaunchKernelGGL((k_grad_accum<CPLX, LPR, UPL, false, true>), gd, bd, 0, st, nbr, b, row_ptr, col_ind,
                               U, V, U2, V2, R, coef, G);
    } else {
        if (b == 1)
            hipLaunchKernelGGL((k_grad_accum<CPLX, LPR, UPL, true, false>), gd, bd, 0, st, nbr, b, row_ptr, col_ind,
                               U, V, U2, V2, R, coef, G);
        else
            hipLaunchKernelGGL((k_grad_accum<CPLX, LPR, UPL, false, false>), gd, bd, 0, st, nbr, b, row_ptr,
                               col_ind, U, V, U2, V2, R, coef, G);
    }
    return hipGetLastError();
}

// Sample columns a lane for 8 < R <= 32: 1 (one column, LPR >= R lanes a
// row), 2 (half the lanes a row, twice the rows a wave) or 4 (R <= 32: 8 lanes
// a row). Measured on the C4-shaped adjoint (65,536 rows x 13 entries, R = 32
// complex): backward 65.4 / 65.8 ms with one column a lane, 61.8 / 60.9 ms
// with two.
int g_grad_wide = 2;

template <bool CPLX>
hipError_t grad_by_r(int64_t nbr, int b, const int64_t *row_ptr, const int32_t *col_ind, const double *U,
                     const double *V, const double *U2, const double *V2, int R, double coef, double *G,
                     hipStream_t st) {
    if (R <= 8) return go_grad<CPLX, 8, 1>(nbr, b, row_ptr, col_ind, U, V, U2, V2, R, coef, G, st);
    if (g_grad_wide >= 2 && R <= 16) return go_grad<CPLX, 8, 2>(nbr, b, row_ptr, col_ind, U, V, U2, V2, R, coef, G, st);
    if (g_grad_wide == 2 && R <= 32) return go_grad<CPLX, 16, 2>(nbr, b, row_ptr, col_ind, U, V, U2, V2, R, coef, G, st);
    if (g_grad_wide == 4 && R <= 32) return go_grad<CPLX, 8, 4>(nbr, b, row_ptr, col_ind, U, V, U2, V2, R, coef, G, st);
    if (R <= 16) return go_grad<CPLX, 16, 1>(nbr, b, row_ptr, col_ind, U, V, U2, V2, R, coef, G, st);
    if (R <= 32) return go_grad<CPLX, 32, 1>(nbr, b, row_ptr, col_ind, U, V, U2, V2, R, coef, G, st);
    if (R <= 64) return go_grad<CPLX, 64, 1>(nbr, b, row_ptr, col_ind, U, V, U2, V2, R, coef, G, st);
    if (R <= 128) return go_grad<CPLX, 64, 2>(nbr, b, row_ptr, col_ind, U, V, U2, V2, R, coef, G, st);
    return go_grad<CPLX, 64, 4>(nbr, b, row_ptr, col_ind, U, V, U2, V2, R, coef, G, st);
}

hipError_t launch_grad_accum(int64_t nbr, int b, int64_t nnzb, const int64_t *row_ptr,
                             const int32_t *col_ind, const double *U, const double *V, int R,
                             bool cplx, double coef, double *G, hipStream_t st) {
    return launch_grad_accum2(nbr, b, nnzb, row_ptr, col_ind, U, V, nullptr, nullptr, R, cplx, coef, G, st);
}

hipError_t launch_grad_accum2(int64_t nbr, int b, int64_t nnzb, const int64_t *row_ptr,
                              const int32_t *col_ind, const double *U, const double *V, const double *U2,
                              const double *V2, int R, bool cplx, double coef, double *G, hipStream_t st) {
    (void)nnzb;
    if (R < 1 || R > 256) return hipErrorInvalidValue;
    if ((U2 == nullptr) != (V2 == nullptr)) return hipErrorInvalidValue;
    return cplx ? grad_by_r<true>(nbr, b, row_ptr, col_ind, U, V, U2, V2, R, coef, G, st)
                : grad_by_r<false>(nbr, b, row_ptr, col_ind, U, V, U2, V2, R, coef, G, st);
}

hipError_t launch_zero(double *y, int64_t n, hipStream_t st) {
    hipLaunchKernelGGL(k_zero, dim3(grid_for(n)), dim3(256), 0, st, y, n);
    return hipGetLastError();
}

hipError_t launch_widen_cols(const float *in, int64_t ld_in, double *out, int64_t rows, int cols, hipStream_t st) {
    hipLaunchKernelGGL(k_widen_cols, dim3(grid_for(rows * cols)), dim3(256), 0, st, in, ld_in, out, rows, cols);
    return hipGetLastError();
}

hipError_t launch_axpby(double *y, const double *x, double alpha, double beta, int64_t n, hipStream_t st) {
    hipLaunchKernelGGL(k_axpby, dim3(grid_for(n)), dim3(256), 0, st, y, x, alpha, beta, n);
    return hipGetLastError();
}

void set_grad_wide(int cols) { g_grad_wide = cols == 1 || cols == 4 ? cols : 2; }

} // namespace rg
