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

#include <algorithm>
#include <numeric>
#include <type_traits>

namespace rg {
namespace {

constexpr int kWaves = 4;   // waves per workgroup
constexpr int kMaxGrid = 2048;

// XCD-contiguous row ranges in the MFMA SpMM (1) or plain round-robin (0).
int g_xcd_map = 1;
// Upper bound on the 16-column tiles one wave owns in the MFMA SpMM.
int g_mfma_ntcap = 16;
// Register prefetch depths of the MFMA SpMM: FP64 plain products 3 (2 when
// set to 2; steps always 2: their epilogue needs the registers), FP32 3 or 4.
// Measured at the diamond 8^3 sandwich, R = 64 complex: FP64 plain 105.6 us
// at 3 against 107.4 at 2, FP64 step 114.9 at 2 against 116.4 at 3; FP32
// plain 54.0 and step 62.2 at 3, 70.8 at 4.
int g_mfma_pd64 = 3;
int g_mfma_pd32 = 3;
// Union-staged path for dense blocks (when the matrix carries group lists).
int g_union_stage = 0;
// LDS-DMA staging depth P of the MFMA SpMM (k_spmm_dma); 0 (the default)
// keeps the register kernel k_spmm_mfma everywhere. With P > 0 the DMA kernel
// runs the FP32 launches with a recurrence epilogue. Since k_spmm_mfma
// prefetches in registers it is the faster of the two at the diamond 8^3
// sandwich (FP32 step 62 us against 81 us); the DMA kernel stays selectable.
int g_dma_depth = 0;
// LDS bytes one union-staged workgroup stages per chunk (two groups per CU).
constexpr size_t kGrpLdsBudget = 72 * 1024;

// Tiles per wave of the MFMA path for W real columns (0: not eligible).
__host__ __device__ inline int mfma_nt(int W) {
    if (W % 32 != 0) return 0;
    if (W % 256 == 0) return 16;
    if (W % 128 == 0) return 8;
    if (W % 64 == 0) return 4;
    return 2;
}

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
                    const MT a = blk[k * bs + r]; // blocks are stored column-major on the device
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
                } else if (mode == EPI_ADJ) {
                    y = sub(affine(y, e.alpha, e.beta, XS[idx]), XP[idx]);
                    y = affine(y, 1.0, e.cm, static_cast<const VT *>(e.qm)[idx]);
                    y = affine(y, 1.0, e.cz, static_cast<const VT *>(e.qz)[idx]);
                    y = affine(y, 1.0, e.cp, static_cast<const VT *>(e.qp)[idx]);
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


// ------------------------------------------------------------ MFMA SpMM
// Dense-block path (b >= 8, real matrix): the block row's stored blocks are
// concatenated along K, so a row of 17 blocks of 13 x 13 is one 13 x 221
// panel times a gathered 221 x W panel of X, tiled as 16 x 16 x 4 MFMAs:
// only the 13 -> 16 row padding is wasted (81 % of issued MFMA work is
// useful), K is padded once per block row (221 -> 224).
//
// Operand layouts of v_mfma_{f64,f32}_16x16x4 (lane l, i = l % 16, g = l / 16):
//   A[i][k = g], B[k = g][j = i]; accumulator element v holds D[4 g + v][i]
//   for f32 and D[4 v + g][i] for f64 (both measured on gfx950).
// N is relabelled so that lane i owns NT *contiguous* real columns
// col0 + i NT + t (tile t): X, Y and the recurrence operands move as wide
// vector loads/stores, and the (re, im) halves of a complex column sit in
// the same lane (tiles 2u, 2u + 1), so the complex dots need no shuffles.
// Blocks are stored column-major on the device, so the 13 rows of one A
// column are contiguous and the A fragment load is coalesced.
template <typename S> struct Acc4;
template <> struct Acc4<double> { typedef double T __attribute__((ext_vector_type(4))); };
template <> struct Acc4<float> { typedef float T __attribute__((ext_vector_type(4))); };

__device__ __forceinline__ Acc4<double>::T mfma16(double a, double b, Acc4<double>::T c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ Acc4<float>::T mfma16(float a, float b, Acc4<float>::T c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// NT contiguous elements; NT * sizeof(S) is a multiple of 8 bytes.
template <typename S, int NT>
__device__ __forceinline__ void load_row(const S *__restrict__ p, S (&v)[NT]) {
    if constexpr ((NT * sizeof(S)) % 16 == 0) {
        const float4 *q = reinterpret_cast<const float4 *>(p);
        constexpr int NQ = NT * sizeof(S) / 16;
#pragma unroll
        for (int t = 0; t < NQ; ++t) {
            const float4 w = q[t];
            __builtin_memcpy(reinterpret_cast<char *>(v) + 16 * t, &w, 16);
        }
    } else {
#pragma unroll
        for (int t = 0; t < NT; ++t) v[t] = p[t];
    }
}
template <typename S, int NT>
__device__ __forceinline__ void store_row(S *__restrict__ p, const S (&v)[NT]) {
    if constexpr ((NT * sizeof(S)) % 16 == 0) {
        float4 *q = reinterpret_cast<float4 *>(p);
        constexpr int NQ = NT * sizeof(S) / 16;
#pragma unroll
        for (int t = 0; t < NQ; ++t) {
            float4 w;
            __builtin_memcpy(&w, reinterpret_cast<const char *>(v) + 16 * t, 16);
            q[t] = w;
        }
    } else {
#pragma unroll
        for (int t = 0; t < NT; ++t) p[t] = v[t];
    }
}

// Epilogue of one finished 16-row tile of the MFMA SpMMs: the rows r < b of
// block row br, columns col0 + i NT + t, from the accumulators; recurrence,
// dots and store as EpiArgs says.
template <typename S, int NT, int MODE, bool CV>
__device__ __forceinline__ void mfma_epilogue(const typename Acc4<S>::T (&acc)[NT], int64_t br,
                                              int b, int W, int col0, int i, int g,
                                              const EpiArgs &e,
                                              double (&dacc)[epi_ndot(MODE) > 0 ? epi_ndot(MODE) : 1][NT],
                                              S *__restrict__ Y) {
    const S *__restrict__ XS = static_cast<const S *>(e.xs);
    const S *__restrict__ XP = static_cast<const S *>(e.xp);
    const S *__restrict__ PSI = static_cast<const S *>(e.psi);
    // one output row per pass keeps the epilogue's registers low (unrolling
    // it costs occupancy and time)
#pragma unroll 1
    for (int v = 0; v < 4; ++v) {
        // accumulator element v holds row 4 v + g (f64) or 4 g + v (f32)
        const int r = sizeof(S) == 8 ? 4 * v + g : 4 * g + v;
        if (r >= b) continue;
        const int64_t base = (br * b + r) * (int64_t)W + col0 + i * NT;
        S y[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            // element v of acc[t], selected without dynamic indexing
            y[t] = v == 0 ? acc[t][0] : v == 1 ? acc[t][1] : v == 2 ? acc[t][2] : acc[t][3];
        }
        if constexpr (MODE != EPI_PLAIN) {
            S xs[NT];
            load_row<S, NT>(XS + base, xs);
#pragma unroll
            for (int t = 0; t < NT; ++t)
                y[t] = (S)__builtin_fma(e.alpha, (double)y[t], e.beta * (double)xs[t]);
            if constexpr (MODE == EPI_STEP || MODE == EPI_PLAINSTEP || MODE == EPI_ADJ) {
                S xp[NT];
                load_row<S, NT>(XP + base, xp);
#pragma unroll
                for (int t = 0; t < NT; ++t) y[t] = y[t] - xp[t];
            }
            if constexpr (MODE == EPI_ADJ) {
                const S *qs[3] = {static_cast<const S *>(e.qm), static_cast<const S *>(e.qz),
                                  static_cast<const S *>(e.qp)};
                const double cs[3] = {e.cm, e.cz, e.cp};
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    S qv[NT];
                    load_row<S, NT>(qs[q] + base, qv);
#pragma unroll
                    for (int t = 0; t < NT; ++t)
                        y[t] = (S)__builtin_fma(cs[q], (double)qv[t], (double)y[t]);
                }
            }
            // dots (FP64): conj(a) b per complex column (t, t+1) or per
            // real column; a self-dot is real and keeps no imaginary part
            auto dot = [&](int d, const S (&a)[NT], const S (&bb)[NT], bool self) {
                if constexpr (CV) {
#pragma unroll
                    for (int t = 0; t < NT; t += 2) {
                        const double ar = a[t], ai = a[t + 1], br_ = bb[t], bi = bb[t + 1];
                        dacc[d][t] = __builtin_fma(ar, br_, __builtin_fma(ai, bi, dacc[d][t]));
                        if (!self)
                            dacc[d][t + 1] = __builtin_fma(ar, bi, __builtin_fma(-ai, br_, dacc[d][t + 1]));
                    }
                } else {
#pragma unroll
                    for (int t = 0; t < NT; ++t)
                        dacc[d][t] = __builtin_fma((double)a[t], (double)bb[t], dacc[d][t]);
                }
            };
            if constexpr (MODE == EPI_FIRST) {
                dot(0, xs, xs, true);
                dot(1, xs, y, false);
                dot(2, y, y, true);
            } else if constexpr (MODE == EPI_STEP) {
                dot(0, y, y, true);
                dot(1, y, xs, false);
            } else if constexpr (MODE == EPI_PLAINSTEP) {
                S ps[NT];
                load_row<S, NT>(PSI + base, ps);
                dot(0, ps, y, false);
            }
        }
        store_row<S, NT>(Y + base, y);
    }
}

// Per-workgroup dot partials of the MFMA SpMMs: the lane groups of every
// column are folded, then the waves sharing a column chunk, into
// e.partials[(blockIdx.x * NDOT + d) * R + u].
template <typename S, int NT, int MODE, bool CV>
__device__ __forceinline__ void mfma_dot_partials(
    double (&dacc)[epi_ndot(MODE) > 0 ? epi_ndot(MODE) : 1][NT], int wave, int i, int g,
    int nchunks, int R, const EpiArgs &e) {
    constexpr int NDOT = epi_ndot(MODE);
    constexpr int CW = 16 * NT;
    if constexpr (NDOT > 0) {
        // fold the four lane groups (rows of the 16-row tile) of every column
#pragma unroll
        for (int d = 0; d < NDOT; ++d)
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                dacc[d][t] += __shfl_xor(dacc[d][t], 16);
                dacc[d][t] += __shfl_xor(dacc[d][t], 32);
            }
        // per wave: CW real columns -> CW/2 complex (or CW real) units
        constexpr int UPC = CV ? CW / 2 : CW;
        __shared__ double2 sm[kWaves][UPC];
        for (int d = 0; d < NDOT; ++d) {
            if (g == 0) {
#pragma unroll
                for (int t = 0; t < NT; t += (CV ? 2 : 1)) {
                    const int uu = CV ? (i * NT + t) / 2 : i * NT + t;
                    sm[wave][uu] = CV ? double2{dacc[d][t], dacc[d][t + 1]} : double2{dacc[d][t], 0.0};
                }
            }
            __syncthreads();
            for (int u = threadIdx.x; u < R; u += blockDim.x) {
                const int ch = (u * (CV ? 2 : 1)) / CW;
                const int uu = u - ch * UPC;
                double2 acc2 = double2{0.0, 0.0};
                for (int w = ch; w < kWaves; w += nchunks) {
                    acc2.x += sm[w][uu].x;
                    acc2.y += sm[w][uu].y;
                }
                e.partials[((int64_t)blockIdx.x * NDOT + d) * R + u] = acc2;
            }
            __syncthreads();
        }
    }
}

template <typename S, int NT, int MODE, bool CV, int PD>
__global__ __launch_bounds__(256) void k_spmm_mfma(int64_t nbr, int b,
                                                   const int64_t *__restrict__ row_ptr,
                                                   const int32_t *__restrict__ xrow,
                                                   const S *__restrict__ vals,
                                                   const S *__restrict__ X, S *__restrict__ Y,
                                                   int W, int R, int xcd_map, EpiArgs e) {
    using A4 = typename Acc4<S>::T;
    constexpr int NDOT = MODE == EPI_FIRST ? 3 : MODE == EPI_STEP ? 2 : MODE == EPI_PLAINSTEP ? 1 : 0;
    constexpr int CW = 16 * NT; // real columns per chunk

    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int i = lane & 15;
    const int g = lane >> 4;
    const int nchunks = W / CW; // divides kWaves: a wave keeps chunk wave % nchunks
    const int col0 = (wave % nchunks) * CW;
    // Workgroups are dispatched round-robin over the 8 XCDs; with xcd_map each
    // XCD walks its own contiguous eighth of the block rows, so the rows in
    // flight on one XCD (and their gathered neighbours) stay close together
    // in its L2.
    int64_t r_lo = 0, r_hi = nbr;
    int64_t kq = blockIdx.x, nk = gridDim.x;
    if (xcd_map) {
        const int x = blockIdx.x & 7;
        kq = blockIdx.x >> 3;
        nk = gridDim.x >> 3;
        r_lo = nbr * x / 8;
        r_hi = nbr * (x + 1) / 8;
    }
    const int64_t total = (r_hi - r_lo) * nchunks;

    double dacc[NDOT > 0 ? NDOT : 1][NT];
#pragma unroll
    for (int d = 0; d < (NDOT > 0 ? NDOT : 1); ++d)
#pragma unroll
        for (int t = 0; t < NT; ++t) dacc[d][t] = 0.0;

    for (int64_t it = kq * kWaves + wave; it < total; it += nk * kWaves) {
        const int64_t br = r_lo + it / nchunks;
        const int64_t s0 = row_ptr[br];
        const int nblk = (int)(row_ptr[br + 1] - s0);
        const int nsteps = (nblk * b + 3) >> 2;

        A4 acc[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[t] = A4{0, 0, 0, 0};

        // Step st covers K columns 4 st .. 4 st + 3 of the row's column-major
        // b x K panel (element (i, k) at s0 b^2 + k b + i); lane group g takes
        // column k = 4 st + g, its X row from the gather table (wave-uniform
        // scalar loads, so no vector load sits between the steps' loads). Two
        // register buffers: step st + 1's operands are in flight while step
        // st's MFMAs run. Past the row's K the table still names a valid row
        // and the panel read stays inside the padded value array; a = 0 there.
        // Lanes outside the panel (i >= b, or k >= K) load their A element from
        // row_ptr[0], which is 0 (checked at upload): the 8 zero bytes read as
        // +0.0, so no select on loaded data holds up the next loads.
        const int K = nblk * b;
        const S *__restrict__ panel = vals + s0 * (int64_t)b * b;
        const S *__restrict__ zero = reinterpret_cast<const S *>(row_ptr);
        int fst = 0; // next step to fetch
        int64_t kb = s0 * b;
        int32_t nx0 = xrow[kb], nx1 = xrow[kb + 1], nx2 = xrow[kb + 2], nx3 = xrow[kb + 3];
        auto fetch = [&](S &a, S (&bv)[NT]) {
            const int64_t xr = g == 0 ? nx0 : g == 1 ? nx1 : g == 2 ? nx2 : nx3;
            const int k = 4 * fst + g;
            a = __builtin_nontemporal_load((i < b && k < K) ? panel + (int64_t)k * b + i : zero);
            load_row<S, NT>(X + xr * W + col0 + i * NT, bv);
            // the next step's table entries (in range: 32 spare entries)
            kb += 4;
            nx0 = xrow[kb];
            nx1 = xrow[kb + 1];
            nx2 = xrow[kb + 2];
            nx3 = xrow[kb + 3];
            ++fst;
        };
        auto mma = [&](S a, const S (&bv)[NT]) {
#pragma unroll
            for (int t = 0; t < NT; ++t) acc[t] = mfma16(a, bv[t], acc[t]);
        };
        // PD register buffers, PD - 1 steps in flight. The fetches are
        // unconditional (up to 2 PD - 2 steps past the row: their A lanes read
        // the zero word and their X rows are valid), so the wait before each
        // MFMA group counts only the later buffers' loads on every path.
        if (nsteps > 0) {
            S ab[PD], bb[PD][NT];
#pragma unroll
            for (int q = 0; q < PD - 1; ++q) fetch(ab[q], bb[q]);
            for (int st = 0; st < nsteps; st += PD) {
#pragma unroll
                for (int q = 0; q < PD; ++q) {
                    fetch(ab[(q + PD - 1) % PD], bb[(q + PD - 1) % PD]);
                    if (st + q < nsteps) mma(ab[q], bb[q]);
                }
            }
        }

        mfma_epilogue<S, NT, MODE, CV>(acc, br, b, W, col0, i, g, e, dacc, Y);
    }

    mfma_dot_partials<S, NT, MODE, CV>(dacc, wave, i, g, nchunks, R, e);
}

// ------------------------------------------------ DMA-staged MFMA SpMM
// The arithmetic of k_spmm_mfma, with the operands of every 4-wide K step
// (the four gathered X row chunks and the wave's 16 x 4 A fragment) moved
// global -> LDS by the LDS-DMA path (global_load_lds, no VGPRs held while in
// flight), P - 1 steps ahead of the MFMAs that consume them. The stream of
// steps runs across the wave's block rows, so the next row's first steps are
// in flight during a row's epilogue. Each wave owns P step buffers of its own
// LDS; only vmcnt orders a buffer's DMA before its reads, and a buffer is
// refilled one step after it was read (the read waits on lgkmcnt), so no
// barrier is needed.

// s_waitcnt vmcnt(VM), then N16 16-byte B chunks (at xaddr + 256 q) and the A
// element (at aaddr, AB bytes) from LDS; the LDS reads complete (lgkmcnt(0))
// before the asm ends, so its outputs are ready when it returns.
using F4 = float __attribute__((ext_vector_type(4)));
template <int VM, int N16, int AB> struct DmaRead;
template <int VM> struct DmaRead<VM, 1, 8> {
    __device__ static __forceinline__ void run(unsigned xaddr, unsigned aaddr, F4 (&o)[1], double &a) {
        asm volatile("s_waitcnt vmcnt(%3)\n\t"
                     "ds_read_b128 %0, %2\n\t"
                     "ds_read_b64 %1, %4\n\t"
                     "s_waitcnt lgkmcnt(0)"
                     : "=&v"(o[0]), "=&v"(a)
                     : "v"(xaddr), "i"(VM), "v"(aaddr)
                     : "memory");
    }
};
template <int VM> struct DmaRead<VM, 2, 8> {
    __device__ static __forceinline__ void run(unsigned xaddr, unsigned aaddr, F4 (&o)[2], double &a) {
        asm volatile("s_waitcnt vmcnt(%4)\n\t"
                     "ds_read_b128 %0, %3\n\t"
                     "ds_read_b128 %1, %3 offset:256\n\t"
                     "ds_read_b64 %2, %5\n\t"
                     "s_waitcnt lgkmcnt(0)"
                     : "=&v"(o[0]), "=&v"(o[1]), "=&v"(a)
                     : "v"(xaddr), "i"(VM), "v"(aaddr)
                     : "memory");
    }
};
template <int VM> struct DmaRead<VM, 4, 8> {
    __device__ static __forceinline__ void run(unsigned xaddr, unsigned aaddr, F4 (&o)[4], double &a) {
        asm volatile("s_waitcnt vmcnt(%6)\n\t"
                     "ds_read_b128 %0, %5\n\t"
                     "ds_read_b128 %1, %5 offset:256\n\t"
                     "ds_read_b128 %2, %5 offset:512\n\t"
                     "ds_read_b128 %3, %5 offset:768\n\t"
                     "ds_read_b64 %4, %7\n\t"
                     "s_waitcnt lgkmcnt(0)"
                     : "=&v"(o[0]), "=&v"(o[1]), "=&v"(o[2]), "=&v"(o[3]), "=&v"(a)
                     : "v"(xaddr), "i"(VM), "v"(aaddr)
                     : "memory");
    }
};
template <int VM> struct DmaRead<VM, 8, 8> {
    __device__ static __forceinline__ void run(unsigned xaddr, unsigned aaddr, F4 (&o)[8], double &a) {
        asm volatile("s_waitcnt vmcnt(%10)\n\t"
                     "ds_read_b128 %0, %9\n\t"
                     "ds_read_b128 %1, %9 offset:256\n\t"
                     "ds_read_b128 %2, %9 offset:512\n\t"
                     "ds_read_b128 %3, %9 offset:768\n\t"
                     "ds_read_b128 %4, %9 offset:1024\n\t"
                     "ds_read_b128 %5, %9 offset:1280\n\t"
                     "ds_read_b128 %6, %9 offset:1536\n\t"
                     "ds_read_b128 %7, %9 offset:1792\n\t"
                     "ds_read_b64 %8, %11\n\t"
                     "s_waitcnt lgkmcnt(0)"
                     : "=&v"(o[0]), "=&v"(o[1]), "=&v"(o[2]), "=&v"(o[3]), "=&v"(o[4]), "=&v"(o[5]), "=&v"(o[6]), "=&v"(o[7]), "=&v"(a)
                     : "v"(xaddr), "i"(VM), "v"(aaddr)
                     : "memory");
    }
};
template <int VM> struct DmaRead<VM, 1, 4> {
    __device__ static __forceinline__ void run(unsigned xaddr, unsigned aaddr, F4 (&o)[1], float &a) {
        asm volatile("s_waitcnt vmcnt(%3)\n\t"
                     "ds_read_b128 %0, %2\n\t"
                     "ds_read_b32 %1, %4\n\t"
                     "s_waitcnt lgkmcnt(0)"
                     : "=&v"(o[0]), "=&v"(a)
                     : "v"(xaddr), "i"(VM), "v"(aaddr)
                     : "memory");
    }
};
template <int VM> struct DmaRead<VM, 2, 4> {
    __device__ static __forceinline__ void run(unsigned xaddr, unsigned aaddr, F4 (&o)[2], float &a) {
        asm volatile("s_waitcnt vmcnt(%4)\n\t"
                     "ds_read_b128 %0, %3\n\t"
                     "ds_read_b128 %1, %3 offset:256\n\t"
                     "ds_read_b32 %2, %5\n\t"
                     "s_waitcnt lgkmcnt(0)"
                     : "=&v"(o[0]), "=&v"(o[1]), "=&v"(a)
                     : "v"(xaddr), "i"(VM), "v"(aaddr)
                     : "memory");
    }
};
template <int VM> struct DmaRead<VM, 4, 4> {
    __device__ static __forceinline__ void run(unsigned xaddr, unsigned aaddr, F4 (&o)[4], float &a) {
        asm volatile("s_waitcnt vmcnt(%6)\n\t"
                     "ds_read_b128 %0, %5\n\t"
                     "ds_read_b128 %1, %5 offset:256\n\t"
                     "ds_read_b128 %2, %5 offset:512\n\t"
                     "ds_read_b128 %3, %5 offset:768\n\t"
                     "ds_read_b32 %4, %7\n\t"
                     "s_waitcnt lgkmcnt(0)"
                     : "=&v"(o[0]), "=&v"(o[1]), "=&v"(o[2]), "=&v"(o[3]), "=&v"(a)
                     : "v"(xaddr), "i"(VM), "v"(aaddr)
                     : "memory");
    }
};
template <int VM> struct DmaRead<VM, 8, 4> {
    __device__ static __forceinline__ void run(unsigned xaddr, unsigned aaddr, F4 (&o)[8], float &a) {
        asm volatile("s_waitcnt vmcnt(%10)\n\t"
                     "ds_read_b128 %0, %9\n\t"
                     "ds_read_b128 %1, %9 offset:256\n\t"
                     "ds_read_b128 %2, %9 offset:512\n\t"
                     "ds_read_b128 %3, %9 offset:768\n\t"
                     "ds_read_b128 %4, %9 offset:1024\n\t"
                     "ds_read_b128 %5, %9 offset:1280\n\t"
                     "ds_read_b128 %6, %9 offset:1536\n\t"
                     "ds_read_b128 %7, %9 offset:1792\n\t"
                     "ds_read_b32 %8, %11\n\t"
                     "s_waitcnt lgkmcnt(0)"
                     : "=&v"(o[0]), "=&v"(o[1]), "=&v"(o[2]), "=&v"(o[3]), "=&v"(o[4]), "=&v"(o[5]), "=&v"(o[6]), "=&v"(o[7]), "=&v"(a)
                     : "v"(xaddr), "i"(VM), "v"(aaddr)
                     : "memory");
    }
};

// 32-bit LDS byte address of a pointer into __shared__ memory
__device__ __forceinline__ unsigned lds_addr(const void *p) {
    return (unsigned)(uintptr_t)((const __attribute__((address_space(3))) char *)p);
}
#define RG_LDS(p) ((__attribute__((address_space(3))) void *)(p))

template <typename S, int NT>
__host__ __device__ constexpr int dma_step_bytes() {
    return 4 * 16 * NT * (int)sizeof(S) + 64 * (int)sizeof(S);
}

template <typename S, int NT, int MODE, bool CV, int P>
__global__ __launch_bounds__(256) void k_spmm_dma(int64_t nbr, int b,
                                                  const int64_t *__restrict__ row_ptr,
                                                  const int32_t *__restrict__ xrow,
                                                  const S *__restrict__ vals,
                                                  const S *__restrict__ X, S *__restrict__ Y,
                                                  int W, int R, int xcd_map, EpiArgs e) {
    using A4 = typename Acc4<S>::T;
    constexpr int NDOT = epi_ndot(MODE);
    constexpr int CW = 16 * NT;
    constexpr int ROWB = CW * (int)sizeof(S); // bytes of one gathered X row chunk
    constexpr int XB = 4 * ROWB;              // X bytes per step
    static_assert(XB % 1024 == 0, "one X DMA instruction moves 1 KB");
    constexpr int NI = XB / 1024;             // X DMA instructions per step
    constexpr int NA = (int)sizeof(S) / 4;    // A DMA instructions per step (256 B each)
    constexpr int STB = dma_step_bytes<S, NT>();
    constexpr int VM = NI + NA;               // vmcnt increments per step
    static_assert(VM * (P - 1) <= 63, "vmcnt holds 6 bits");
    constexpr int N16 = NT * (int)sizeof(S) / 16;
    static_assert(N16 >= 1, "B chunk of at least 16 bytes");
    extern __shared__ __attribute__((aligned(16))) unsigned char dma_lds[];

    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int i = lane & 15;
    const int g = lane >> 4;
    const int nchunks = W / CW; // 1, 2 or 4 (divides kWaves): a wave keeps chunk wave % nchunks
    const int cshift = nchunks == 4 ? 2 : nchunks == 2 ? 1 : 0;
    const int col0 = (wave % nchunks) * CW;
    int64_t r_lo = 0, r_hi = nbr;
    int64_t kq = blockIdx.x, nk = gridDim.x;
    if (xcd_map) { // as k_spmm_mfma: each XCD walks its own eighth of the rows
        const int x = blockIdx.x & 7;
        kq = blockIdx.x >> 3;
        nk = gridDim.x >> 3;
        r_lo = nbr * x / 8;
        r_hi = nbr * (x + 1) / 8;
    }
    const int64_t total = (r_hi - r_lo) * nchunks;
    const int64_t wstride = nk * kWaves;
    unsigned char *wbuf = dma_lds + wave * (P * STB);
    const unsigned wbuf_addr = lds_addr(wbuf);
    const char *Xc = reinterpret_cast<const char *>(X) + (size_t)col0 * sizeof(S);
    const char *Vc = reinterpret_cast<const char *>(vals);
    // LDS order of a row chunk: 16-byte chunk c = i N16 + q of the row (lane
    // i's q-th chunk) sits at position q 16 + i, so each of a lane group's N16
    // ds_read_b128 touches 256 contiguous bytes (no bank conflicts). The DMA
    // lane filling position pos of the step's X region therefore fetches row
    // pos / (16 N16), chunk ((pp & 15) N16 + pp / 16) of it (pp: position in
    // the row).

    // ---- issue side: the wave's K steps in consumption order. Within a
    // block row the blocks are consecutive and column-major, so the row's A
    // panel is one column-major b x K matrix: step st's four columns are the
    // 4 b contiguous elements from (s0 b + 4 st) b. The step's X rows come
    // from the gather table at s0 b + 4 st.
    int64_t iit = kq * kWaves + wave; // work item whose steps are being issued
    int64_t is0 = 0;                  // its first block
    int insteps = 0, ist = 0;
    auto open_row = [&]() { // from iit, the next work item with at least one step
        while (iit < total) {
            const int64_t br = r_lo + (iit >> cshift);
            is0 = row_ptr[br];
            insteps = ((int)(row_ptr[br + 1] - is0) * b + 3) >> 2;
            if (insteps > 0) break;
            iit += wstride;
        }
        ist = 0;
    };
    open_row();
    // DMAs the next step into buffer `slot`; false once the stream is done
    auto issue = [&](int slot) -> bool {
        if (iit >= total) return false;
        unsigned char *buf = wbuf + slot * STB;
        const int64_t kb = is0 * b + 4 * ist;
        int64_t xr[4];
#pragma unroll
        for (int kr = 0; kr < 4; ++kr) xr[kr] = (int64_t)xrow[kb + kr] * W;
#pragma unroll
        for (int j = 0; j < NI; ++j) {
            // instruction j fills positions [64 j, 64 j + 64) of the step
            const int pos = j * 64 + lane;
            const int kr = pos / (16 * N16);
            const int pp = pos - kr * (16 * N16);
            const int src = ((pp & 15) * N16 + (pp >> 4)) * 16;
            int64_t r;
            if constexpr (ROWB >= 1024) r = xr[(j * 1024) / ROWB];
            else r = kr == 0 ? xr[0] : kr == 1 ? xr[1] : kr == 2 ? xr[2] : xr[3];
            __builtin_amdgcn_global_load_lds(Xc + r * (int64_t)sizeof(S) + src,
                                             RG_LDS(buf + j * 1024), 16, 0, 0);
        }
        const char *asrc = Vc + (kb * b) * (int64_t)sizeof(S) + 4 * lane;
#pragma unroll
        for (int q = 0; q < NA; ++q)
            __builtin_amdgcn_global_load_lds(asrc + 256 * q, RG_LDS(buf + XB + 256 * q), 4, 0, 0);
        if (++ist == insteps) {
            iit += wstride;
            open_row();
        }
        return true;
    };

    double dacc[NDOT > 0 ? NDOT : 1][NT];
#pragma unroll
    for (int d = 0; d < (NDOT > 0 ? NDOT : 1); ++d)
#pragma unroll
        for (int t = 0; t < NT; ++t) dacc[d][t] = 0.0;

#pragma unroll
    for (int q = 0; q < P - 1; ++q) (void)issue(q);
    int cs = 0; // step buffer consumed next
    const unsigned xoff = g * ROWB + i * 16;
    for (int64_t it = kq * kWaves + wave; it < total; it += wstride) {
        const int64_t br = r_lo + (it >> cshift);
        const int K = (int)(row_ptr[br + 1] - row_ptr[br]) * b;
        const int nsteps = (K + 3) >> 2;
        A4 acc[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[t] = A4{0, 0, 0, 0};
        for (int st = 0; st < nsteps; ++st) {
            // refill the buffer the previous step read, then wait for this one
            const bool more = issue(cs == 0 ? P - 1 : cs - 1);
            const unsigned sb = wbuf_addr + cs * STB;
            const unsigned xa = sb + xoff;
            const unsigned aa = sb + XB + (g * b + i) * (int)sizeof(S);
            F4 o[N16];
            S a;
            if (more) DmaRead<VM * (P - 1), N16, (int)sizeof(S)>::run(xa, aa, o, a);
            else DmaRead<0, N16, (int)sizeof(S)>::run(xa, aa, o, a);
            if (!(i < b && 4 * st + g < K)) a = S(0);
            S bv[NT];
            __builtin_memcpy(bv, o, sizeof(bv));
#pragma unroll
            for (int t = 0; t < NT; ++t) acc[t] = mfma16(a, bv[t], acc[t]);
            cs = cs + 1 == P ? 0 : cs + 1;
        }
        mfma_epilogue<S, NT, MODE, CV>(acc, br, b, W, col0, i, g, e, dacc, Y);
    }
    mfma_dot_partials<S, NT, MODE, CV>(dacc, wave, i, g, nchunks, R, e);
}

template <typename S, int NT, bool CV, int PD>
hipError_t go_mfma_pd(const BsrDev &A, const void *X, void *Y, int W, int R, const EpiArgs &e,
                      int grid, hipStream_t st) {
    const int xcd = g_xcd_map;
    const dim3 gd(grid), bd(64 * kWaves);
    const S *vals = static_cast<const S *>(A.vals);
    const S *x = static_cast<const S *>(X);
    S *y = static_cast<S *>(Y);
    switch (e.mode) {
    case EPI_PLAIN:
        hipLaunchKernelGGL((k_spmm_mfma<S, NT, EPI_PLAIN, CV, PD>), gd, bd, 0, st, A.nbr, (int)A.b,
                           A.row_ptr, A.xrow, vals, x, y, W, R, xcd, e);
        break;
    case EPI_FIRST:
        hipLaunchKernelGGL((k_spmm_mfma<S, NT, EPI_FIRST, CV, PD>), gd, bd, 0, st, A.nbr, (int)A.b,
                           A.row_ptr, A.xrow, vals, x, y, W, R, xcd, e);
        break;
    case EPI_STEP:
        hipLaunchKernelGGL((k_spmm_mfma<S, NT, EPI_STEP, CV, PD>), gd, bd, 0, st, A.nbr, (int)A.b,
                           A.row_ptr, A.xrow, vals, x, y, W, R, xcd, e);
        break;
    case EPI_PLAINSTEP:
        hipLaunchKernelGGL((k_spmm_mfma<S, NT, EPI_PLAINSTEP, CV, PD>), gd, bd, 0, st, A.nbr,
                           (int)A.b, A.row_ptr, A.xrow, vals, x, y, W, R, xcd, e);
        break;
    default:
        hipLaunchKernelGGL((k_spmm_mfma<S, NT, EPI_ADJ, CV, PD>), gd, bd, 0, st, A.nbr,
                           (int)A.b, A.row_ptr, A.xrow, vals, x, y, W, R, xcd, e);
        break;
    }
    return hipGetLastError();
}

template <typename S, int NT, int MODE, bool CV, int P>
hipError_t go_dma_mode(const BsrDev &A, const void *X, void *Y, int W, int R, const EpiArgs &e,
                       int grid, hipStream_t st) {
    constexpr unsigned lds = kWaves * P * dma_step_bytes<S, NT>();
    static bool attr_set = false; // one attribute call per kernel instance
    if (!attr_set) {
        hipError_t he = hipFuncSetAttribute(
            reinterpret_cast<const void *>(&k_spmm_dma<S, NT, MODE, CV, P>),
            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (he != hipSuccess) return he;
        attr_set = true;
    }
    hipLaunchKernelGGL((k_spmm_dma<S, NT, MODE, CV, P>), dim3(grid), dim3(64 * kWaves), lds, st,
                       A.nbr, (int)A.b, A.row_ptr, A.xrow, static_cast<const S *>(A.vals),
                       static_cast<const S *>(X), static_cast<S *>(Y), W, R, g_xcd_map, e);
    return hipGetLastError();
}

template <typename S, int NT, bool CV, int P>
hipError_t go_dma_p(const BsrDev &A, const void *X, void *Y, int W, int R, const EpiArgs &e,
                    int grid, hipStream_t st) {
    switch (e.mode) {
    case EPI_PLAIN: return go_dma_mode<S, NT, EPI_PLAIN, CV, P>(A, X, Y, W, R, e, grid, st);
    case EPI_FIRST: return go_dma_mode<S, NT, EPI_FIRST, CV, P>(A, X, Y, W, R, e, grid, st);
    case EPI_STEP: return go_dma_mode<S, NT, EPI_STEP, CV, P>(A, X, Y, W, R, e, grid, st);
    case EPI_PLAINSTEP: return go_dma_mode<S, NT, EPI_PLAINSTEP, CV, P>(A, X, Y, W, R, e, grid, st);
    default: return go_dma_mode<S, NT, EPI_ADJ, CV, P>(A, X, Y, W, R, e, grid, st);
    }
}

template <typename S, int NT, bool CV>
hipError_t go_dma(const BsrDev &A, const void *X, void *Y, int W, int R, const EpiArgs &e, int grid,
                  hipStream_t st) {
    if (g_dma_depth == 4) return go_dma_p<S, NT, CV, 4>(A, X, Y, W, R, e, grid, st);
    return go_dma_p<S, NT, CV, 2>(A, X, Y, W, R, e, grid, st);
}

// NT * sizeof(S) >= 16: FP64 with NT >= 2, FP32 with NT >= 4
template <typename S, bool CV>
hipError_t dma_by_nt(int nt, const BsrDev &A, const void *X, void *Y, int W, int R,
                     const EpiArgs &e, int grid, hipStream_t st) {
    switch (nt) {
    case 2:
        if constexpr (sizeof(S) == 8) return go_dma<S, 2, CV>(A, X, Y, W, R, e, grid, st);
        return hipErrorInvalidValue;
    case 4: return go_dma<S, 4, CV>(A, X, Y, W, R, e, grid, st);
    case 8: return go_dma<S, 8, CV>(A, X, Y, W, R, e, grid, st);
    default: return go_dma<S, 16, CV>(A, X, Y, W, R, e, grid, st);
    }
}

// register prefetch depth of k_spmm_mfma (g_mfma_pd64, g_mfma_pd32)
template <typename S, int NT, bool CV>
hipError_t go_mfma(const BsrDev &A, const void *X, void *Y, int W, int R, const EpiArgs &e,
                   int grid, hipStream_t st) {
    if constexpr (sizeof(S) == 8) {
        if (g_mfma_pd64 == 3 && e.mode == EPI_PLAIN)
            return go_mfma_pd<S, NT, CV, 3>(A, X, Y, W, R, e, grid, st);
        return go_mfma_pd<S, NT, CV, 2>(A, X, Y, W, R, e, grid, st);
    } else {
        if (g_mfma_pd32 == 4) return go_mfma_pd<S, NT, CV, 4>(A, X, Y, W, R, e, grid, st);
        return go_mfma_pd<S, NT, CV, 3>(A, X, Y, W, R, e, grid, st);
    }
}

template <typename S, bool CV>
hipError_t by_nt(int nt, const BsrDev &A, const void *X, void *Y, int W, int R, const EpiArgs &e,
                 int grid, hipStream_t st) {
    switch (nt) {
    case 2: return go_mfma<S, 2, CV>(A, X, Y, W, R, e, grid, st);
    case 4: return go_mfma<S, 4, CV>(A, X, Y, W, R, e, grid, st);
    case 8: return go_mfma<S, 8, CV>(A, X, Y, W, R, e, grid, st);
    default: return go_mfma<S, 16, CV>(A, X, Y, W, R, e, grid, st);
    }
}

// ---------------------------------------------- union-staged MFMA SpMM
// The 8 block rows of a workgroup (one wave each) gather, between them, far
// fewer distinct X block panels than 8 x 17: consecutive diamond atoms share
// neighbours, so a group of 8 needs ~56 distinct panels instead of 136. The
// group's union list (sorted block columns, host-built once per matrix) is
// streamed through LDS C panels at a time; each wave runs its own blocks that
// fall in the staged chunk (K concatenated within the chunk) with its B
// fragments read from LDS. Two workgroups per CU overlap one group's staging
// with the other's MFMAs. Every wave owns all W columns (NT = W / 16).
// LDS rows are swizzled in 16-byte units so that the 16 lanes reading one K
// row (64 contiguous bytes each) hit distinct banks.
constexpr int kGrpWaves = 8;

template <typename S, int NT>
__device__ __forceinline__ int grp_swz(int u) { // logical -> physical 16-byte unit
    constexpr int NQ = NT * (int)sizeof(S) / 16;
    if constexpr (NQ <= 1) {
        return u;
    } else {
        const int i = u / NQ, q = u - (u / NQ) * NQ;
        return i * NQ + (q ^ ((i * NQ / 16) & (NQ - 1)));
    }
}

template <typename S, int NT, int MODE, bool CV>
__global__ __launch_bounds__(512) void k_spmm_grp(int64_t nbr, int b,
                                                  const int64_t *__restrict__ row_ptr,
                                                  const int32_t *__restrict__ col_ind,
                                                  const S *__restrict__ vals,
                                                  const int64_t *__restrict__ grp_ptr,
                                                  const int32_t *__restrict__ grp_cols,
                                                  const int32_t *__restrict__ pos,
                                                  const S *__restrict__ X, S *__restrict__ Y,
                                                  int R, int C, int xcd_map, EpiArgs e) {
    using A4 = typename Acc4<S>::T;
    constexpr int NDOT = MODE == EPI_FIRST ? 3 : MODE == EPI_STEP ? 2 : MODE == EPI_PLAINSTEP ? 1 : 0;
    constexpr int W = 16 * NT;                 // real columns, all in one wave
    constexpr int UW = W * (int)sizeof(S) / 16; // 16-byte units per staged row
    extern __shared__ float4 lds_units[];
    const S *__restrict__ XS = static_cast<const S *>(e.xs);
    const S *__restrict__ XP = static_cast<const S *>(e.xp);
    const S *__restrict__ PSI = static_cast<const S *>(e.psi);

    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int i = lane & 15;
    const int g = lane >> 4;
    const int64_t ngroups = (nbr + kGrpWaves - 1) / kGrpWaves;
    // XCD x walks the x-th contiguous eighth of the groups
    int64_t g_lo = 0, g_hi = ngroups, kq = blockIdx.x, nk = gridDim.x;
    if (xcd_map) {
        const int64_t per = (ngroups + 7) / 8;
        const int x = blockIdx.x & 7;
        kq = blockIdx.x >> 3;
        nk = gridDim.x >> 3;
        g_lo = x * per;
        g_hi = min(ngroups, (int64_t)(x + 1) * per);
    }

    double dacc[NDOT > 0 ? NDOT : 1][NT];
#pragma unroll
    for (int d = 0; d < (NDOT > 0 ? NDOT : 1); ++d)
#pragma unroll
        for (int t = 0; t < NT; ++t) dacc[d][t] = 0.0;

    for (int64_t grp = g_lo + kq; grp < g_hi; grp += nk) {
        const int64_t br = grp * kGrpWaves + wave;
        const bool has_row = br < nbr;
        const int64_t u0 = grp_ptr[grp];
        const int nu = (int)(grp_ptr[grp + 1] - u0);
        int64_t sp = has_row ? row_ptr[br] : 0;
        const int64_t s1 = has_row ? row_ptr[br + 1] : 0;

        A4 acc[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[t] = A4{0, 0, 0, 0};

        for (int cu = 0; cu < nu; cu += C) {
            const int nc = min(C, nu - cu);
            // stage panels cu .. cu + nc - 1: b rows of W elements each
            const int units = nc * b * UW;
            for (int t = threadIdx.x; t < units; t += blockDim.x) {
                const int rowl = t / UW, uu = t - rowl * UW;
                const int j = rowl / b, kk = rowl - j * b;
                const int64_t xr = (int64_t)grp_cols[u0 + cu + j] * b + kk;
                const float4 v = reinterpret_cast<const float4 *>(X + xr * W)[uu];
                lds_units[rowl * UW + grp_swz<S, NT>(uu)] = v;
            }
            __syncthreads();
            if (has_row) {
                int64_t se = sp;
                while (se < s1 && pos[se] < cu + nc) ++se;
                const int nk_ = (int)(se - sp) * b;
                const int nsteps = (nk_ + 3) >> 2;
                int blk = 0, kk = g;
                while (kk >= b) { kk -= b; ++blk; }
                for (int st = 0; st < nsteps; ++st) {
                    S a = S(0);
                    S bv[NT];
                    if (4 * st + g < nk_) {
                        const int64_t sb = sp + blk;
                        if (i < b) a = __builtin_nontemporal_load(vals + sb * (int64_t)(b * b) + kk * b + i);
                        const int rowl = (pos[sb] - cu) * b + kk;
                        constexpr int NQ = NT * (int)sizeof(S) / 16;
                        const float4 *src = lds_units + rowl * UW;
#pragma unroll
                        for (int q = 0; q < NQ; ++q) {
                            const float4 w = src[grp_swz<S, NT>(i * NQ + q)];
                            __builtin_memcpy(reinterpret_cast<char *>(bv) + 16 * q, &w, 16);
                        }
                    } else {
#pragma unroll
                        for (int t = 0; t < NT; ++t) bv[t] = S(0);
                    }
#pragma unroll
                    for (int t = 0; t < NT; ++t) acc[t] = mfma16(a, bv[t], acc[t]);
                    kk += 4;
                    while (kk >= b) { kk -= b; ++blk; }
                }
                sp = se;
            }
            __syncthreads();
        }

        if (!has_row) continue;
#pragma unroll 1
        for (int v = 0; v < 4; ++v) {
            const int r = sizeof(S) == 8 ? 4 * v + g : 4 * g + v;
            if (r >= b) continue;
            const int64_t base = (br * b + r) * (int64_t)W + i * NT;
            S y[NT];
#pragma unroll
            for (int t = 0; t < NT; ++t)
                y[t] = v == 0 ? acc[t][0] : v == 1 ? acc[t][1] : v == 2 ? acc[t][2] : acc[t][3];
            if constexpr (MODE != EPI_PLAIN) {
                S xs[NT];
                load_row<S, NT>(XS + base, xs);
#pragma unroll
                for (int t = 0; t < NT; ++t)
                    y[t] = (S)__builtin_fma(e.alpha, (double)y[t], e.beta * (double)xs[t]);
                if constexpr (MODE == EPI_STEP || MODE == EPI_PLAINSTEP || MODE == EPI_ADJ) {
                    S xp[NT];
                    load_row<S, NT>(XP + base, xp);
#pragma unroll
                    for (int t = 0; t < NT; ++t) y[t] = y[t] - xp[t];
                }
                if constexpr (MODE == EPI_ADJ) {
                    const S *qs[3] = {static_cast<const S *>(e.qm), static_cast<const S *>(e.qz),
                                      static_cast<const S *>(e.qp)};
                    const double cs[3] = {e.cm, e.cz, e.cp};
#pragma unroll
                    for (int q = 0; q < 3; ++q) {
                        S qv[NT];
                        load_row<S, NT>(qs[q] + base, qv);
#pragma unroll
                        for (int t = 0; t < NT; ++t)
                            y[t] = (S)__builtin_fma(cs[q], (double)qv[t], (double)y[t]);
                    }
                }
                auto dot = [&](int d, const S (&a)[NT], const S (&bb)[NT], bool self) {
                    if constexpr (CV) {
#pragma unroll
                        for (int t = 0; t < NT; t += 2) {
                            const double ar = a[t], ai = a[t + 1], br_ = bb[t], bi = bb[t + 1];
                            dacc[d][t] = __builtin_fma(ar, br_, __builtin_fma(ai, bi, dacc[d][t]));
                            if (!self)
                                dacc[d][t + 1] = __builtin_fma(ar, bi, __builtin_fma(-ai, br_, dacc[d][t + 1]));
                        }
                    } else {
#pragma unroll
                        for (int t = 0; t < NT; ++t)
                            dacc[d][t] = __builtin_fma((double)a[t], (double)bb[t], dacc[d][t]);
                    }
                };
                if constexpr (MODE == EPI_FIRST) {
                    dot(0, xs, xs, true);
                    dot(1, xs, y, false);
                    dot(2, y, y, true);
                } else if constexpr (MODE == EPI_STEP) {
                    dot(0, y, y, true);
                    dot(1, y, xs, false);
                } else if constexpr (MODE == EPI_PLAINSTEP) {
                    S ps[NT];
                    load_row<S, NT>(PSI + base, ps);
                    dot(0, ps, y, false);
                }
            }
            store_row<S, NT>(Y + base, y);
        }
    }

    if constexpr (NDOT > 0) {
#pragma unroll
        for (int d = 0; d < NDOT; ++d)
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                dacc[d][t] += __shfl_xor(dacc[d][t], 16);
                dacc[d][t] += __shfl_xor(dacc[d][t], 32);
            }
        constexpr int UPC = CV ? W / 2 : W;
        double2 *sm = reinterpret_cast<double2 *>(lds_units); // staging area is free now
        for (int d = 0; d < NDOT; ++d) {
            if (g == 0) {
#pragma unroll
                for (int t = 0; t < NT; t += (CV ? 2 : 1)) {
                    const int uu = CV ? (i * NT + t) / 2 : i * NT + t;
                    sm[wave * UPC + uu] = CV ? double2{dacc[d][t], dacc[d][t + 1]} : double2{dacc[d][t], 0.0};
                }
            }
            __syncthreads();
            for (int u = threadIdx.x; u < R; u += blockDim.x) {
                double2 a2 = double2{0.0, 0.0};
                for (int w = 0; w < kGrpWaves; ++w) {
                    a2.x += sm[w * UPC + u].x;
                    a2.y += sm[w * UPC + u].y;
                }
                e.partials[((int64_t)blockIdx.x * NDOT + d) * R + u] = a2;
            }
            __syncthreads();
        }
    }
}

template <typename S, int NT, bool CV>
hipError_t go_grp(const BsrDev &A, const void *X, void *Y, int R, const EpiArgs &e, int grid,
                  int C, size_t lds, hipStream_t st) {
    const dim3 gd(grid), bd(64 * kGrpWaves);
    const S *vals = static_cast<const S *>(A.vals);
    const S *x = static_cast<const S *>(X);
    S *y = static_cast<S *>(Y);
    const int xcd = g_xcd_map;
#define RG_GRP_LAUNCH(MODE)                                                                        \
    hipLaunchKernelGGL((k_spmm_grp<S, NT, MODE, CV>), gd, bd, lds, st, A.nbr, (int)A.b, A.row_ptr,  \
                       A.col_ind, vals, A.grp_ptr, A.grp_cols, A.pos, x, y, R, C, xcd, e)
    switch (e.mode) {
    case EPI_PLAIN: RG_GRP_LAUNCH(EPI_PLAIN); break;
    case EPI_FIRST: RG_GRP_LAUNCH(EPI_FIRST); break;
    case EPI_STEP: RG_GRP_LAUNCH(EPI_STEP); break;
    case EPI_PLAINSTEP: RG_GRP_LAUNCH(EPI_PLAINSTEP); break;
    default: RG_GRP_LAUNCH(EPI_ADJ); break;
    }
#undef RG_GRP_LAUNCH
    return hipGetLastError();
}

template <typename S, bool CV>
hipError_t grp_by_nt(const BsrDev &A, const void *X, void *Y, int W, int R, const EpiArgs &e,
                     int grid, int C, size_t lds, hipStream_t st) {
    switch (W / 16) {
    case 2: return go_grp<S, 2, CV>(A, X, Y, R, e, grid, C, lds, st);
    case 4: return go_grp<S, 4, CV>(A, X, Y, R, e, grid, C, lds, st);
    case 8: return go_grp<S, 8, CV>(A, X, Y, R, e, grid, C, lds, st);
    default: return go_grp<S, 16, CV>(A, X, Y, R, e, grid, C, lds, st);
    }
}

// ------------------------------------------------------------ reduction
// Neumaier-compensated sum of a and b into (s, c): s + c carries the exact
// sum to twice the working precision (Neumaier 1974, "improved Kahan").
__device__ __forceinline__ void two_sum_acc(double &s, double &c, double x) {
    const double t = s + x;
    c += (fabs(s) >= fabs(x)) ? ((s - t) + x) : ((x - t) + s);
    s = t;
}

__global__ __launch_bounds__(256) void k_reduce(const double2 *__restrict__ partials, int ngroups,
                                                int ndot, int R, double2 *__restrict__ out) {
    const int lane = threadIdx.x & 63;
    const int pair = blockIdx.x * kWaves + (threadIdx.x >> 6);
    const int npairs = ndot * R;
    if (pair >= npairs) return;
    const int d = pair / R;
    const int u = pair - d * R;
    // compensated FP64 sums over the workgroup partials, then a fixed xor tree
    double sx = 0.0, cx = 0.0, sy = 0.0, cy = 0.0;
    // eight loads in flight per lane before they are summed in order
    for (int g0 = lane; g0 < ngroups; g0 += 64 * 8) {
        double2 p[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int gi = g0 + 64 * q;
            p[q] = gi < ngroups ? partials[((int64_t)gi * ndot + d) * R + u] : double2{0.0, 0.0};
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            two_sum_acc(sx, cx, p[q].x);
            two_sum_acc(sy, cy, p[q].y);
        }
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const double ox = __shfl_xor(sx, off), ocx = __shfl_xor(cx, off);
        const double oy = __shfl_xor(sy, off), ocy = __shfl_xor(cy, off);
        cx += ocx;
        cy += ocy;
        two_sum_acc(sx, cx, ox);
        two_sum_acc(sy, cy, oy);
    }
    if (lane == 0) out[pair] = double2{sx + cx, sy + cy};
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

// Row-major -> column-major within every b x b block (real or complex).
__global__ __launch_bounds__(256) void k_block_transpose(const double *__restrict__ in,
                                                         double *__restrict__ out, int64_t nnzb,
                                                         int b, int cplx) {
    const int64_t per = (int64_t)b * b;
    const int64_t n = nnzb * per;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t s = t / per;
        const int e = (int)(t - s * per);
        const int r = e / b, c = e - (e / b) * b; // source (r, c) row-major
        const int64_t dst = s * per + (int64_t)c * b + r;
        if (cplx) {
            out[2 * dst] = in[2 * t];
            out[2 * dst + 1] = in[2 * t + 1];
        } else {
            out[dst] = in[t];
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
    if (A.is_complex && !vm.complex_vec) return hipErrorInvalidValue;
    const int W = R * (vm.complex_vec ? 2 : 1);
    const size_t elem = vm.fp32 ? 4 : 8;
    // each lane owns W/16 columns; they must fill whole 16-byte LDS units
    if (g_union_stage && !A.is_complex && A.b >= 8 && A.grp_ptr &&
        (W == 32 || W == 64 || W == 128 || W == 256) && W * elem >= 256) {
        const size_t panel = (size_t)A.b * W * elem;
        int C = (int)(kGrpLdsBudget / panel);
        if (C > 16) C = 16;
        if (C >= 1) {
            const int64_t ngroups = (A.nbr + kGrpWaves - 1) / kGrpWaves;
            int64_t g = g_xcd_map ? ((ngroups + 7) / 8) * 8 : ngroups;
            if (g > kMaxGrid) g = kMaxGrid; // grid-stride over groups (a multiple of 8)
            {
                const int grid = (int)g;
                if (grid_out) *grid_out = grid;
                const size_t lds = std::max((size_t)C * panel, (size_t)kGrpWaves * (W / (vm.complex_vec ? 2 : 1)) * 16);
                if (vm.fp32)
                    return vm.complex_vec ? grp_by_nt<float, true>(A, X, Y, W, R, e, grid, C, lds, stream)
                                          : grp_by_nt<float, false>(A, X, Y, W, R, e, grid, C, lds, stream);
                return vm.complex_vec ? grp_by_nt<double, true>(A, X, Y, W, R, e, grid, C, lds, stream)
                                      : grp_by_nt<double, false>(A, X, Y, W, R, e, grid, C, lds, stream);
            }
        }
    }
    int nt = mfma_nt(W);
    while (nt > 2 && nt > g_mfma_ntcap) nt /= 2;
    const int nchunks = nt > 0 ? W / (16 * nt) : 0;
    if (!A.is_complex && A.b >= 8 && A.xrow && nt > 0 && kWaves % nchunks == 0) {
        int64_t g = (A.nbr * nchunks + kWaves - 1) / kWaves;
        if (g > kMaxGrid) g = kMaxGrid;
        g = ((g + 7) / 8) * 8; // whole XCD rounds (kMaxGrid is a multiple of 8)
        const int grid = (int)g;
        if (grid_out) *grid_out = grid;
        if (g_dma_depth > 0 && A.xrow && nt * (vm.fp32 ? 4 : 8) >= 16 && vm.fp32 &&
            e.mode != EPI_PLAIN) {
            if (vm.fp32)
                return vm.complex_vec ? dma_by_nt<float, true>(nt, A, X, Y, W, R, e, grid, stream)
                                      : dma_by_nt<float, false>(nt, A, X, Y, W, R, e, grid, stream);
            return vm.complex_vec ? dma_by_nt<double, true>(nt, A, X, Y, W, R, e, grid, stream)
                                  : dma_by_nt<double, false>(nt, A, X, Y, W, R, e, grid, stream);
        }
        if (vm.fp32) {
            return vm.complex_vec ? by_nt<float, true>(nt, A, X, Y, W, R, e, grid, stream)
                                  : by_nt<float, false>(nt, A, X, Y, W, R, e, grid, stream);
        }
        return vm.complex_vec ? by_nt<double, true>(nt, A, X, Y, W, R, e, grid, stream)
                              : by_nt<double, false>(nt, A, X, Y, W, R, e, grid, stream);
    }
    const int grid = spmm_grid(A.nbr, R);
    if (grid_out) *grid_out = grid;
    if (!vm.fp32) {
        if (A.is_complex) return by_block<double, true, true>(A, X, Y, R, e, grid, stream);
        if (vm.complex_vec) return by_block<double, false, true>(A, X, Y, R, e, grid, stream);
        return by_block<double, false, false>(A, X, Y, R, e, grid, stream);
    }
    if (A.is_complex) return by_block<float, true, true>(A, X, Y, R, e, grid, stream);
    if (vm.complex_vec) return by_block<float, false, true>(A, X, Y, R, e, grid, stream);
    return by_block<float, false, false>(A, X, Y, R, e, grid, stream);
}

int spmm_max_grid() { return kMaxGrid; }

void set_xcd_map(int on) { g_xcd_map = on ? 1 : 0; }
void set_mfma_ntcap(int nt) { g_mfma_ntcap = nt; }
void set_union_stage(int on) { g_union_stage = on ? 1 : 0; }
void set_mfma_pd32(int p) { g_mfma_pd32 = p == 4 ? 4 : 3; }
void set_mfma_pd64(int p) { g_mfma_pd64 = p == 2 ? 2 : 3; }
void set_dma_depth(int p) { g_dma_depth = (p == 0 || p == 4) ? p : 2; }
int union_group_rows() { return kGrpWaves; }

hipError_t launch_block_transpose(const double *in, double *out, int64_t nnzb, int b, bool cplx,
                                  hipStream_t stream) {
    hipLaunchKernelGGL(k_block_transpose, dim3(grid_1d(nnzb * b * b)), dim3(256), 0, stream, in,
                       out, nnzb, b, cplx ? 1 : 0);
    return hipGetLastError();
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
