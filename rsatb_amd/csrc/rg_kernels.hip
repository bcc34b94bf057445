// gfx950 kernels of the KPM moment loop.
//
// k_spmm: block-row-gather BSR SpMM Y = epi(A X) over a block of R sample
// columns, fused with the Chebyshev recurrence and the per-column moment
// dots (the three passes rsatb runs separately: mult_dense, axpby, dot;
// rsatb/calc/cheby.hpp:555-606, rsatb/calc/dos.hpp:150-173).
//
// Mapping. A wavefront (64 lanes) owns G = 64/LPR block rows; within a row,
// lane q owns sample columns q*UPL .. q*UPL+UPL-1. With LPR = 64 (R > 32) a
// wave owns exactly one block row, so the block row, its CSR range and every
// matrix entry are wave-uniform and are fetched by scalar loads and fed to
// the FMAs as SGPR operands: the matrix is streamed from HBM once per apply
// and never occupies VGPRs. Each lane keeps b x UPL accumulators; per matrix
// entry it issues UPL (real vectors) or 2 UPL (complex vectors) FMAs, with
// one coalesced vector load per block column (64 lanes x UPL x 8/16 B
// contiguous). With R <= 32 (LPR = R rounded up to 8/16/32) several block
// rows share a wave and entries come through vector loads.
//
// Dots. Every lane accumulates its columns' dot products in FP64 over all
// rows it visits (grid-stride over block rows), then the workgroup reduces
// over sub-rows (xor shuffles) and waves (LDS) and writes one partial per
// column; k_reduce sums the partials of all workgroups in a fixed order, so
// results are deterministic run to run.
#include "rg_internal.h"

#include <type_traits>

namespace rg {
namespace {

constexpr int kWaves = 4;   // waves per workgroup
constexpr int kMaxGrid = 2048;

template <typename S> struct V2;
template <> struct V2<double> { using T = double2; };
template <> struct V2<float> { using T = float2; };

// ---------------------------------------------------------------- algebra
__device__ __forceinline__ void mac(double &acc, double a, double x) { acc = __builtin_fma(a, x, acc); }
__device__ __forceinline__ void mac(float &acc, float a, float x) { acc = __builtin_fmaf(a, x, acc); }
__device__ __forceinline__ void mac(double2 &acc, double a, double2 x) {
    acc.x = __builtin_fma(a, x.x, acc.x);
    acc.y = __builtin_fma(a, x.y, acc.y);
}
__device__ __forceinline__ void mac(float2 &acc, float a, float2 x) {
    acc.x = __builtin_fmaf(a, x.x, acc.x);
    acc.y = __builtin_fmaf(a, x.y, acc.y);
}
__device__ __forceinline__ void mac(double2 &acc, double2 a, double2 x) {
    acc.x = __builtin_fma(a.x, x.x, __builtin_fma(-a.y, x.y, acc.x));
    acc.y = __builtin_fma(a.x, x.y, __builtin_fma(a.y, x.x, acc.y));
}
__device__ __forceinline__ void mac(float2 &acc, float2 a, float2 x) {
    acc.x = __builtin_fmaf(a.x, x.x, __builtin_fmaf(-a.y, x.y, acc.x));
    acc.y = __builtin_fmaf(a.x, x.y, __builtin_fmaf(a.y, x.x, acc.y));
}

template <typename T> __device__ __forceinline__ T vzero() { return T{}; }

// y = alpha v + beta s (- p)
__device__ __forceinline__ double affine(double v, double alpha, double beta, double s) {
    return __builtin_fma(alpha, v, beta * s);
}
__device__ __forceinline__ double2 affine(double2 v, double alpha, double beta, double2 s) {
    return double2{__builtin_fma(alpha, v.x, beta * s.x), __builtin_fma(alpha, v.y, beta * s.y)};
}
__device__ __forceinline__ float affine(float v, double alpha, double beta, float s) {
    return (float)__builtin_fma(alpha, (double)v, beta * (double)s);
}
__device__ __forceinline__ float2 affine(float2 v, double alpha, double beta, float2 s) {
    return float2{(float)__builtin_fma(alpha, (double)v.x, beta * (double)s.x),
                  (float)__builtin_fma(alpha, (double)v.y, beta * (double)s.y)};
}
__device__ __forceinline__ double sub(double a, double b) { return a - b; }
__device__ __forceinline__ float sub(float a, float b) { return a - b; }
__device__ __forceinline__ double2 sub(double2 a, double2 b) { return double2{a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ float2 sub(float2 a, float2 b) { return float2{a.x - b.x, a.y - b.y}; }

// d += conj(a) b in FP64. Products of two floats are exact in FP64.
__device__ __forceinline__ void dotacc(double2 &d, double a, double b) { d.x = __builtin_fma(a, b, d.x); }
__device__ __forceinline__ void dotacc(double2 &d, float a, float b) { d.x = __builtin_fma((double)a, (double)b, d.x); }
__device__ __forceinline__ void dotacc(double2 &d, double2 a, double2 b) {
    d.x = __builtin_fma(a.x, b.x, __builtin_fma(a.y, b.y, d.x));
    d.y = __builtin_fma(a.x, b.y, __builtin_fma(-a.y, b.x, d.y));
}
__device__ __forceinline__ void dotacc(double2 &d, float2 a, float2 b) {
    dotacc(d, double2{(double)a.x, (double)a.y}, double2{(double)b.x, (double)b.y});
}

__device__ __forceinline__ int dev_ndot(int mode) {
    return mode == EPI_FIRST ? 3 : mode == EPI_STEP ? 2 : mode == EPI_PLAINSTEP ? 1 : 0;
}

// ------------------------------------------------------------------ SpMM
template <typename S, bool CM, bool CV, int BS, int LPR, int UPL>
__global__ __launch_bounds__(256) void k_spmm(int64_t nbr, int bs_rt,
                                              const int64_t *__restrict__ row_ptr,
                                              const int32_t *__restrict__ col_ind,
                                              const void *__restrict__ vals_v,
                                              const void *__restrict__ x_v, void *__restrict__ y_v,
                                              int R, EpiArgs e) {
    using VT = std::conditional_t<CV, typename V2<S>::T, S>;
    using MT = std::conditional_t<CM, typename V2<S>::T, S>;
    constexpr int G = 64 / LPR;
    constexpr int MAXB = BS > 0 ? BS : 16;
    constexpr int COLS = LPR * UPL; // columns one sub-row covers

    const MT *__restrict__ vals = static_cast<const MT *>(vals_v);
    const VT *__restrict__ X = static_cast<const VT *>(x_v);
    VT *__restrict__ Y = static_cast<VT *>(y_v);
    const VT *__restrict__ XS = static_cast<const VT *>(e.xs);
    const VT *__restrict__ XP = static_cast<const VT *>(e.xp);
    const VT *__restrict__ PSI = static_cast<const VT *>(e.psi);

    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int g = (G > 1) ? (lane / LPR) : 0;
    const int q = (G > 1) ? (lane % LPR) : lane;
    const int bs = BS > 0 ? BS : bs_rt;
    const int mode = e.mode;
    const int ndot = dev_ndot(mode);

    double2 dacc[3][UPL];
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int j = 0; j < UPL; ++j) dacc[d][j] = double2{0.0, 0.0};

    const int64_t rows_per_wg = (int64_t)kWaves * G;
    for (int64_t base = (int64_t)blockIdx.x * rows_per_wg; base < nbr;
         base += (int64_t)gridDim.x * rows_per_wg) {
        const int64_t br = base + (int64_t)wave * G + g;
        if (br >= nbr) continue;

        VT acc[MAXB][UPL];
#pragma unroll
        for (int r = 0; r < MAXB; ++r)
#pragma unroll
            for (int j = 0; j < UPL; ++j) acc[r][j] = vzero<VT>();

        const int64_t s0 = row_ptr[br];
        const int64_t s1 = row_ptr[br + 1];
        for (int64_t s = s0; s < s1; ++s) {
            const int64_t bc = col_ind[s];
            const MT *__restrict__ blk = vals + s * (int64_t)(bs * bs);
            const VT *__restrict__ xb = X + bc * (int64_t)bs * R;
#pragma unroll
            for (int k = 0; k < MAXB; ++k) {
                if (BS == 0 && k >= bs) continue;
                VT xk[UPL];
#pragma unroll
                for (int j = 0; j < UPL; ++j) {
                    const int u = q * UPL + j;
                    xk[j] = (u < R) ? xb[(int64_t)k * R + u] : vzero<VT>();
                }
#pragma unroll
                for (int r = 0; r < MAXB; ++r) {
                    if (BS == 0 && r >= bs) continue;
                    const MT a = blk[r * bs + k];
#pragma unroll
                    for (int j = 0; j < UPL; ++j) mac(acc[r][j], a, xk[j]);
                }
            }
        }

#pragma unroll
        for (int r = 0; r < MAXB; ++r) {
            if (BS == 0 && r >= bs) continue;
            const int64_t row = br * bs + r;
#pragma unroll
            for (int j = 0; j < UPL; ++j) {
                const int u = q * UPL + j;
                if (u >= R) continue;
                const int64_t idx = row * R + u;
                VT y = acc[r][j];
                if (mode == EPI_FIRST) {
                    const VT xs = XS[idx];
                    y = affine(y, e.alpha, e.beta, xs);
                    dotacc(dacc[0][j], xs, xs);
                    dotacc(dacc[1][j], xs, y);
                    dotacc(dacc[2][j], y, y);
                } else if (mode == EPI_STEP) {
                    const VT xs = XS[idx];
                    y = sub(affine(y, e.alpha, e.beta, xs), XP[idx]);
                    dotacc(dacc[0][j], y, y);
                    dotacc(dacc[1][j], y, xs);
                } else if (mode == EPI_PLAINSTEP) {
                    y = sub(affine(y, e.alpha, e.beta, XS[idx]), XP[idx]);
                    dotacc(dacc[0][j], PSI[idx], y);
                }
                Y[idx] = y;
            }
        }
    }

    if (ndot == 0) return;

    // sub-rows of one wave hold the same columns: fold them (xor tree)
    if (G > 1) {
#pragma unroll
        for (int off = LPR; off < 64; off <<= 1)
#pragma unroll
            for (int d = 0; d < 3; ++d)
#pragma unroll
                for (int j = 0; j < UPL; ++j) {
                    dacc[d][j].x += __shfl_xor(dacc[d][j].x, off);
                    dacc[d][j].y += __shfl_xor(dacc[d][j].y, off);
                }
    }
    __shared__ double2 sm[kWaves][COLS];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (d >= ndot) break;
        if (g == 0) {
#pragma unroll
            for (int j = 0; j < UPL; ++j) sm[wave][q * UPL + j] = dacc[d][j];
        }
        __syncthreads();
        for (int u = threadIdx.x; u < COLS && u < R; u += blockDim.x) {
            double2 acc2 = sm[0][u];
#pragma unroll
            for (int w = 1; w < kWaves; ++w) {
                acc2.x += sm[w][u].x;
                acc2.y += sm[w][u].y;
            }
            e.partials[((int64_t)blockIdx.x * ndot + d) * R + u] = acc2;
        }
        __syncthreads();
    }
}

int lpr_for(int R) { return R <= 8 ? 8 : R <= 16 ? 16 : R <= 32 ? 32 : 64; }
int upl_for(int R) { return R <= 64 ? 1 : R <= 128 ? 2 : 4; }

template <typename S, bool CM, bool CV, int BS, int LPR, int UPL>
hipError_t go(const BsrDev &A, const void *X, void *Y, int R, const EpiArgs &e, int grid,
              hipStream_t st) {
    hipLaunchKernelGGL((k_spmm<S, CM, CV, BS, LPR, UPL>), dim3(grid), dim3(64 * kWaves), 0, st,
                       A.nbr, (int)A.b, A.row_ptr, A.col_ind, A.vals, X, Y, R, e);
    return hipGetLastError();
}

template <typename S, bool CM, bool CV, int BS>
hipError_t by_layout(const BsrDev &A, const void *X, void *Y, int R, const EpiArgs &e, int grid,
                     hipStream_t st) {
    switch (lpr_for(R)) {
    case 8: return go<S, CM, CV, BS, 8, 1>(A, X, Y, R, e, grid, st);
    case 16: return go<S, CM, CV, BS, 16, 1>(A, X, Y, R, e, grid, st);
    case 32: return go<S, CM, CV, BS, 32, 1>(A, X, Y, R, e, grid, st);
    default: break;
    }
    switch (upl_for(R)) {
    case 1: return go<S, CM, CV, BS, 64, 1>(A, X, Y, R, e, grid, st);
    case 2: return go<S, CM, CV, BS, 64, 2>(A, X, Y, R, e, grid, st);
    default: return go<S, CM, CV, BS, 64, 4>(A, X, Y, R, e, grid, st);
    }
}

template <typename S, bool CM, bool CV>
hipError_t by_block(const BsrDev &A, const void *X, void *Y, int R, const EpiArgs &e, int grid,
                    hipStream_t st) {
    if (A.b == 1) return by_layout<S, CM, CV, 1>(A, X, Y, R, e, grid, st);
    if (A.b == 13) return by_layout<S, CM, CV, 13>(A, X, Y, R, e, grid, st);
    return by_layout<S, CM, CV, 0>(A, X, Y, R, e, grid, st);
}

// ------------------------------------------------------------ reduction
__global__ __launch_bounds__(256) void k_reduce(const double2 *__restrict__ partials, int ngroups,
                                                int ndot, int R, double2 *__restrict__ out) {
    const int lane = threadIdx.x & 63;
    const int pair = blockIdx.x * kWaves + (threadIdx.x >> 6);
    const int npairs = ndot * R;
    if (pair >= npairs) return;
    const int d = pair / R;
    const int u = pair - d * R;
    double sx = 0.0, sy = 0.0;
    for (int gi = lane; gi < ngroups; gi += 64) {
        const double2 p = partials[((int64_t)gi * ndot + d) * R + u];
        sx += p.x;
        sy += p.y;
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        sx += __shfl_xor(sx, off);
        sy += __shfl_xor(sy, off);
    }
    if (lane == 0) out[pair] = double2{sx, sy};
}

// --------------------------------------------------------------- samples
__device__ __forceinline__ uint64_t mix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

__global__ __launch_bounds__(256) void k_samples(int kind, uint64_t seed, int64_t rows, int R,
                                                 double scale, int complex_vec, void *out) {
    const int64_t n = rows * (int64_t)R;
    const uint64_t ms = mix64(seed);
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = t / R;
        const int64_t c = t - i * R;
        const uint64_t h = mix64(mix64(ms ^ (uint64_t)i) ^ (uint64_t)c);
        if (kind == 0) { // RademacherSign (rsatb/calc/rs_ldos.hpp:256-264)
            const double v = (h & 1ull) ? scale : -scale;
            if (complex_vec) static_cast<double2 *>(out)[t] = double2{v, 0.0};
            else static_cast<double *>(out)[t] = v;
        } else { // phase 2 pi u, u = top 53 bits * 2^-53
            const double u = (double)(h >> 11) * (1.0 / 9007199254740992.0);
            double sn, cs;
            sincos(6.283185307179586476925286766559 * u, &sn, &cs);
            static_cast<double2 *>(out)[t] = double2{cs * scale, sn * scale};
        }
    }
}

__global__ __launch_bounds__(256) void k_f64_to_f32(const double *__restrict__ in,
                                                    float *__restrict__ out, int64_t n) {
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (int64_t)gridDim.x * blockDim.x)
        out[t] = (float)in[t];
}

__global__ __launch_bounds__(256) void k_stage(const double *__restrict__ in, int in_complex,
                                               void *__restrict__ out, int out_complex,
                                               int out_fp32, int64_t n) {
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (int64_t)gridDim.x * blockDim.x) {
        const double re = in_complex ? in[2 * t] : in[t];
        const double im = in_complex ? in[2 * t + 1] : 0.0;
        if (out_fp32) {
            float *o = static_cast<float *>(out);
            if (out_complex) {
                o[2 * t] = (float)re;
                o[2 * t + 1] = (float)im;
            } else {
                o[t] = (float)re;
            }
        } else {
            double *o = static_cast<double *>(out);
            if (out_complex) {
                o[2 * t] = re;
                o[2 * t + 1] = im;
            } else {
                o[t] = re;
            }
        }
    }
}

int grid_1d(int64_t n) {
    int64_t g = (n + 255) / 256;
    if (g > 8192) g = 8192;
    if (g < 1) g = 1;
    return (int)g;
}

} // namespace

int spmm_grid(int64_t nbr, int R) {
    const int G = 64 / lpr_for(R);
    const int64_t rows_per_wg = (int64_t)kWaves * G;
    int64_t g = (nbr + rows_per_wg - 1) / rows_per_wg;
    if (g > kMaxGrid) g = kMaxGrid;
    if (g < 1) g = 1;
    return (int)g;
}

hipError_t launch_spmm(const BsrDev &A, VecMode vm, const void *X, void *Y, int R,
                       const EpiArgs &e, hipStream_t stream, int *grid_out) {
    if (R < 1 || R > 256 || A.b < 1 || A.b > 16) return hipErrorInvalidValue;
    const int grid = spmm_grid(A.nbr, R);
    if (grid_out) *grid_out = grid;
    if (A.is_complex && !vm.complex_vec) return hipErrorInvalidValue;
    if (!vm.fp32) {
        if (A.is_complex) return by_block<double, true, true>(A, X, Y, R, e, grid, stream);
        if (vm.complex_vec) return by_block<double, false, true>(A, X, Y, R, e, grid, stream);
        return by_block<double, false, false>(A, X, Y, R, e, grid, stream);
    }
    if (A.is_complex) return by_block<float, true, true>(A, X, Y, R, e, grid, stream);
    if (vm.complex_vec) return by_block<float, false, true>(A, X, Y, R, e, grid, stream);
    return by_block<float, false, false>(A, X, Y, R, e, grid, stream);
}

hipError_t launch_reduce(const double2 *partials, int ngroups, int ndot, int R, double2 *out,
                         hipStream_t stream) {
    const int npairs = ndot * R;
    const int grid = (npairs + kWaves - 1) / kWaves;
    hipLaunchKernelGGL(k_reduce, dim3(grid), dim3(64 * kWaves), 0, stream, partials, ngroups, ndot,
                       R, out);
    return hipGetLastError();
}

hipError_t launch_samples(int kind, uint64_t seed, int64_t rows, int R, double scale,
                          bool complex_vec, void *out, hipStream_t stream) {
    hipLaunchKernelGGL(k_samples, dim3(grid_1d(rows * R)), dim3(256), 0, stream, kind, seed, rows,
                       R, scale, complex_vec ? 1 : 0, out);
    return hipGetLastError();
}

hipError_t launch_f64_to_f32(const double *in, float *out, int64_t n, hipStream_t stream) {
    hipLaunchKernelGGL(k_f64_to_f32, dim3(grid_1d(n)), dim3(256), 0, stream, in, out, n);
    return hipGetLastError();
}

hipError_t launch_stage(const double *in, bool in_complex, void *out, bool out_complex,
                        bool out_fp32, int64_t n, hipStream_t stream) {
    hipLaunchKernelGGL(k_stage, dim3(grid_1d(n)), dim3(256), 0, stream, in, in_complex ? 1 : 0,
                       out, out_complex ? 1 : 0, out_fp32 ? 1 : 0, n);
    return hipGetLastError();
}

} // namespace rg
