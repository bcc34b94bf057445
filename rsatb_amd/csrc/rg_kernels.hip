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

#include <atomic>

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
// Cap on the MFMA SpMM grid (workgroups, at most kMaxGrid); 0: two
// workgroups per compute unit. A grid-stride loop over fewer workgroups keeps
// the block rows in flight on each XCD closer together than the 2048-workgroup
// grid that covers every row once. Measured at the diamond 24^3 sandwich
// (R = 128 complex; the FP64 kernels hold one workgroup per CU, the FP32 plain
// product three), 2048 / resident / 512 / 256 workgroups: FP64 plain product
// 4994 / 4807 / 4638 / 4808 us, FP32 plain 2886 / 2870 / 2560 / 2709 us;
// diamond 8^3 FP64 plain 93.9 / 92.8 / 87.5 / 102.4 us.
int g_mfma_grid_cap = 0;
// b = 1 SpMM over real vectors with 2 columns per lane: 1 always, 0 never,
// 2 by precision and row length.
int g_b1_wide = 2;
// The VALU SpMM computes the entry range of uniform rows (1) or loads it (0).
int g_uniform_rows = 1;
// Bytes of a row one lane owns in the b = 1 slab SpMM (16, 32 or 64; 0: the
// column-per-lane layout of k_spmm), and the fewest workgroups a slab launch
// may have (below it the rows go to k_spmm, whose layouts spread them over
// more waves). Measured on graphene p_z, R = 16 real, FP64 / FP32 step at
// 1,048,576 rows: k_spmm 134.1 / 69.7 us, slabs of 16 bytes 92.3 / 42.3 us
// (0.63 / 0.72 of the HBM peak), 32 bytes 99.2 / 45.0, 64 bytes 117.7 / 65.8;
// the 16,384-row graphene run (FP64 / FP32) 2.90 / 2.73 ms with k_spmm,
// 2.61 / 2.86 ms with 16-byte slabs: the FP32 slab launch there has 256
// workgroups, one wave per SIMD.
// Automatic (-1): 16 bytes, 32 for complex FP64 vectors (two columns a lane:
// the C4-shaped adjoint, 65,536 rows x 13 entries, R = 32 complex, forward
// 25.1 -> 23.7 ms and backward 70.4 -> 69.5 ms against 16 bytes).
int g_b1v_bytes = -1;
int g_b1v_min_grid = 512;
// Register prefetch depths of the MFMA SpMM: FP64 plain products 3 (2 when
// set to 2; steps always 2: their epilogue needs the registers), FP32 3 or 4.
// Measured at the diamond 8^3 sandwich, R = 64 complex: FP64 plain 105.6 us
// at 3 against 107.4 at 2, FP64 step 114.9 at 2 against 116.4 at 3; FP32
// plain 54.0 and step 62.2 at 3, 70.8 at 4.
int g_mfma_pd64 = 3;
int g_mfma_pd32 = 3;
// The same for sixteen tiles per wave over complex vectors, [FP64, FP32]
// x [plain product, step]: at sixteen tiles the FP64 kernels hold one wave
// per SIMD (128 accumulator registers), so a wave hides its gathers' latency
// only with the steps it has in flight. Measured at the diamond 24^3
// sandwich, R = 128 complex, depth 2 / 3 / 4 / 5 / 6: FP64 plain 4727 /
// 4522 / 4583 / 4845 / 4926 us, FP64 step (2 / 3 / 4 / 5 / 6) 5348 / 5474 /
// 5510 / 5701 / 5752 us; FP32 (plain, step) at 2: 2329, 2699 us, at 3: 2414,
// 2749 us, at 4: 2450, 3421 us. Deeper buffers do not pay: with the rows in
// flight held to one region per XCD the limit is the L2, not the latency.
int g_pd16[2][2] = {{3, 2}, {2, 2}};

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

// s += x with the rounding error carried in c (TwoSum, Knuth 1969: exact
// for any s and x; s + c is the sum to about twice the working precision).
__device__ __forceinline__ void two_sum_add(double &s, double &c, double x) {
    const double t = s + x;
    const double bp = t - s;
    c += (s - (t - bp)) + (x - bp);
    s = t;
}

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
__global__ __launch_bounds__(256) void k_spmm(int64_t nbr, int bs_rt, int ulen,
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

        // uniform rows: the range follows from br, and the column index
        // loads wait on nothing (one dependent load fewer per row)
        const int64_t s0 = ulen > 0 ? br * ulen : row_ptr[br];
        const int64_t s1 = ulen > 0 ? s0 + ulen : row_ptr[br + 1];
        int64_t s = s0;
        if constexpr (BS == 1) {
            // Scalar entries: the gather of each entry's X row waits on its
            // column index, so U entries' indices and values are loaded
            // first and their U row loads issued back to back (the
            // accumulation order is unchanged). U = 4 measured best: 4-entry
            // graphene rows take it whole (FP64 step 161 -> 137 us at
            // 1,048,576 rows), 22-entry rows gain over U = 8 (74 -> 67 us),
            // and complex vectors stay at 8 waves per SIMD.
            constexpr int U = 4;
            for (; s + U <= s1; s += U) {
                int64_t bc[U];
                MT a[U];
#pragma unroll
                for (int t = 0; t < U; ++t) {
                    bc[t] = col_ind[s + t];
                    a[t] = vals[s + t];
                }
                VT xk[U][UPL];
#pragma unroll
                for (int t = 0; t < U; ++t)
#pragma unroll
                    for (int j = 0; j < UPL; ++j) {
                        const int u = q * UPL + j;
                        xk[t][j] = (u < R) ? X[bc[t] * R + u] : vzero<VT>();
                    }
#pragma unroll
                for (int t = 0; t < U; ++t)
#pragma unroll
                    for (int j = 0; j < UPL; ++j) mac(acc[0][j], a[t], xk[t][j]);
            }
        }
        for (; s < s1; ++s) {
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

// The wide b = 1 layout (half the lanes per row, two columns per lane) for
// scalar entries and real vectors, 8 < R <= 32: FP32, or rows of >= 8
// entries (g_b1_wide 2), always (1) or never (0).
bool b1_wide(const BsrDev &A, bool fp32, bool cv, int R) {
    if (A.b != 1 || cv || R <= 8 || R > 32) return false;
    return g_b1_wide == 1 || (g_b1_wide == 2 && (fp32 || A.nnzb >= 8 * A.nbr));
}

// Lanes per block row of the VALU SpMM launch for this operand.
int b1v_lpr(const BsrDev &A, bool fp32, bool cv, int R);
int valu_lpr(const BsrDev &A, bool fp32, bool cv, int R) {
    if (const int l = b1v_lpr(A, fp32, cv, R)) return l;
    if (b1_wide(A, fp32, cv, R)) return R <= 16 ? 8 : 16;
    return lpr_for(R);
}

template <typename S, bool CM, bool CV, int BS, int LPR, int UPL>
hipError_t go(const BsrDev &A, const void *X, void *Y, int R, const EpiArgs &e, int grid,
              hipStream_t st) {
    hipLaunchKernelGGL((k_spmm<S, CM, CV, BS, LPR, UPL>), dim3(grid), dim3(64 * kWaves), 0, st,
                       A.nbr, (int)A.b, g_uniform_rows ? A.ulen : 0, A.row_ptr, A.col_ind, A.vals, X, Y, R, e);
    return hipGetLastError();
}

template <typename S, bool CM, bool CV, int BS>
hipError_t by_layout(const BsrDev &A, const void *X, void *Y, int R, const EpiArgs &e, int grid,
                     hipStream_t st) {
    if constexpr (BS == 1 && !CV) {
        // Scalar entries, real vectors: half the lanes per row and two
        // columns per lane, twice the rows (and independent gathers) per
        // wave. Measured on graphene p_z with R = 16: FP32 step 103 -> 76 us
        // at 1,048,576 rows of 4 entries, FP64 86 -> 75 us at 190,528 rows of
        // 22 entries, but FP64 163 -> 193 us at 4 entries a row; complex
        // vectors (R = 32) slower. Automatic: FP32, or rows of >= 8 entries.
        const bool wide = b1_wide(A, sizeof(S) == 4, CV, R);
        if (wide && R <= 16) return go<S, CM, CV, BS, 8, 2>(A, X, Y, R, e, grid, st);
        if (wide) return go<S, CM, CV, BS, 16, 2>(A, X, Y, R, e, grid, st);
    }
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

// ------------------------------------------------------- b = 1 slab SpMM
// Scalar entries, real matrix, R = LPR * CPL exactly: lane q of a row's LPR
// lanes owns the contiguous slab of columns q * CPL .. q * CPL + CPL - 1 and
// moves it as 16-byte loads and stores (CPL values of 8 or 16 bytes: 16, 32
// or 64 bytes a lane). Fewer lanes per row put more rows, and more
// independent gathers, in one wave than the column-per-lane layout of
// k_spmm; every product and epilogue value is formed by the same operations
// in the same order as there (the dots differ only in their summation
// order). Vector operands must be 16-byte aligned.
template <typename S, bool CV, int LPR, int CPL>
__global__ __launch_bounds__(256) void k_spmm_b1v(int64_t nbr, int ulen, const int64_t *__restrict__ row_ptr,
                                                  const int32_t *__restrict__ col_ind,
                                                  const S *__restrict__ vals, const S *__restrict__ X,
                                                  S *__restrict__ Y, EpiArgs e) {
    using VT = std::conditional_t<CV, typename V2<S>::T, S>;
    constexpr int G = 64 / LPR;
    constexpr int NS = CPL * (CV ? 2 : 1);    // scalars a lane owns in a row
    constexpr int EPC = 16 / (int)sizeof(S);  // scalars a 16-byte chunk holds
    constexpr int NCH = NS / EPC;             // chunks a lane moves per row
    static_assert(NS % EPC == 0 && 64 % LPR == 0, "slab of whole 16-byte chunks");
    typedef S C16 __attribute__((ext_vector_type(EPC)));
    constexpr int64_t W = (int64_t)LPR * NS;  // scalars a row holds

    const S *__restrict__ XS = static_cast<const S *>(e.xs);
    const S *__restrict__ XP = static_cast<const S *>(e.xp);
    const S *__restrict__ PSI = static_cast<const S *>(e.psi);

    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int g = lane / LPR;
    const int q = lane % LPR;
    const int mode = e.mode;
    const int ndot = dev_ndot(mode);

    auto ld = [](const S *p, S (&v)[NS]) {
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            const C16 t = *reinterpret_cast<const C16 *>(p + c * EPC);
#pragma unroll
            for (int k = 0; k < EPC; ++k) v[c * EPC + k] = t[k];
        }
    };
    auto vt = [](const S (&v)[NS], int j) -> VT {
        if constexpr (CV) return VT{v[2 * j], v[2 * j + 1]};
        else return v[j];
    };

    double2 dacc[3][CPL];
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int j = 0; j < CPL; ++j) dacc[d][j] = double2{0.0, 0.0};

    const int64_t rows_per_wg = (int64_t)kWaves * G;
    for (int64_t base = (int64_t)blockIdx.x * rows_per_wg; base < nbr; base += (int64_t)gridDim.x * rows_per_wg) {
        const int64_t br = base + (int64_t)wave * G + g;
        if (br >= nbr) continue;
        S acc[NS];
#pragma unroll
        for (int i = 0; i < NS; ++i) acc[i] = S(0);
        const int64_t s0 = ulen > 0 ? br * ulen : row_ptr[br];
        const int64_t s1 = ulen > 0 ? s0 + ulen : row_ptr[br + 1];
        const int64_t lo = (int64_t)q * NS;
        int64_t s = s0;
        constexpr int U = 4;
        for (; s + U <= s1; s += U) {
            int64_t bc[U];
            S a[U];
#pragma unroll
            for (int t = 0; t < U; ++t) {
                bc[t] = col_ind[s + t];
                a[t] = vals[s + t];
            }
            S xk[U][NS];
#pragma unroll
            for (int t = 0; t < U; ++t) ld(X + bc[t] * W + lo, xk[t]);
#pragma unroll
            for (int t = 0; t < U; ++t)
#pragma unroll
                for (int i = 0; i < NS; ++i) mac(acc[i], a[t], xk[t][i]);
        }
        for (; s < s1; ++s) {
            const int64_t bc = col_ind[s];
            const S a = vals[s];
            S xk[NS];
            ld(X + bc * W + lo, xk);
#pragma unroll
            for (int i = 0; i < NS; ++i) mac(acc[i], a, xk[i]);
        }

        const int64_t idx = br * W + lo;
        S y[NS];
#pragma unroll
        for (int i = 0; i < NS; ++i) y[i] = acc[i];
        if (mode == EPI_FIRST) {
            S xs[NS];
            ld(XS + idx, xs);
#pragma unroll
            for (int i = 0; i < NS; ++i) y[i] = affine(y[i], e.alpha, e.beta, xs[i]);
#pragma unroll
            for (int j = 0; j < CPL; ++j) {
                dotacc(dacc[0][j], vt(xs, j), vt(xs, j));
                dotacc(dacc[1][j], vt(xs, j), vt(y, j));
                dotacc(dacc[2][j], vt(y, j), vt(y, j));
            }
        } else if (mode == EPI_STEP) {
            S xs[NS], xp[NS];
            ld(XS + idx, xs);
            ld(XP + idx, xp);
#pragma unroll
            for (int i = 0; i < NS; ++i) y[i] = sub(affine(y[i], e.alpha, e.beta, xs[i]), xp[i]);
#pragma unroll
            for (int j = 0; j < CPL; ++j) {
                dotacc(dacc[0][j], vt(y, j), vt(y, j));
                dotacc(dacc[1][j], vt(y, j), vt(xs, j));
            }
        } else if (mode == EPI_PLAINSTEP) {
            S xs[NS], xp[NS], ps[NS];
            ld(XS + idx, xs);
            ld(XP + idx, xp);
            ld(PSI + idx, ps);
#pragma unroll
            for (int i = 0; i < NS; ++i) y[i] = sub(affine(y[i], e.alpha, e.beta, xs[i]), xp[i]);
#pragma unroll
            for (int j = 0; j < CPL; ++j) dotacc(dacc[0][j], vt(ps, j), vt(y, j));
        } else if (mode == EPI_ADJ) {
            S xs[NS], xp[NS], t[NS];
            ld(XS + idx, xs);
            ld(XP + idx, xp);
#pragma unroll
            for (int i = 0; i < NS; ++i) y[i] = sub(affine(y[i], e.alpha, e.beta, xs[i]), xp[i]);
            ld(static_cast<const S *>(e.qm) + idx, t);
#pragma unroll
            for (int i = 0; i < NS; ++i) y[i] = affine(y[i], 1.0, e.cm, t[i]);
            ld(static_cast<const S *>(e.qz) + idx, t);
#pragma unroll
            for (int i = 0; i < NS; ++i) y[i] = affine(y[i], 1.0, e.cz, t[i]);
            ld(static_cast<const S *>(e.qp) + idx, t);
#pragma unroll
            for (int i = 0; i < NS; ++i) y[i] = affine(y[i], 1.0, e.cp, t[i]);
        }
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            C16 t;
#pragma unroll
            for (int k = 0; k < EPC; ++k) t[k] = y[c * EPC + k];
            *reinterpret_cast<C16 *>(Y + idx + c * EPC) = t;
        }
    }

    if (ndot == 0) return;
    // the rows of one wave hold the same columns: fold them (xor tree)
    if (G > 1) {
#pragma unroll
        for (int off = LPR; off < 64; off <<= 1)
#pragma unroll
            for (int d = 0; d < 3; ++d)
#pragma unroll
                for (int j = 0; j < CPL; ++j) {
                    dacc[d][j].x += __shfl_xor(dacc[d][j].x, off);
                    dacc[d][j].y += __shfl_xor(dacc[d][j].y, off);
                }
    }
    constexpr int COLS = LPR * CPL;
    __shared__ double2 sm[kWaves][COLS];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (d >= ndot) break;
        if (g == 0) {
#pragma unroll
            for (int j = 0; j < CPL; ++j) sm[wave][q * CPL + j] = dacc[d][j];
        }
        __syncthreads();
        for (int u = threadIdx.x; u < COLS; u += blockDim.x) {
            double2 acc2 = sm[0][u];
#pragma unroll
            for (int w = 1; w < kWaves; ++w) {
                acc2.x += sm[w][u].x;
                acc2.y += sm[w][u].y;
            }
            e.partials[((int64_t)blockIdx.x * ndot + d) * COLS + u] = acc2;
        }
        __syncthreads();
    }
}

// Lanes per row of the slab layout for this launch shape, or 0 when it does
// not apply: b = 1, real matrix, R of 16 or 32, g_b1v_bytes a lane.
int b1v_lpr(const BsrDev &A, bool fp32, bool cv, int R) {
    if (A.b != 1 || A.is_complex || g_b1v_bytes == 0 || (R != 16 && R != 32)) return 0;
    const int vt = (fp32 ? 4 : 8) * (cv ? 2 : 1);
    const int bytes = g_b1v_bytes > 0 ? g_b1v_bytes : (vt == 16 ? 32 : 16);
    const int cpl = bytes / vt;
    if (cpl < 1 || R % cpl != 0) return 0;
    const int lpr = R / cpl;
    const int64_t rows_per_wg = (int64_t)kWaves * (64 / lpr);
    if ((A.nbr + rows_per_wg - 1) / rows_per_wg < g_b1v_min_grid) return 0;
    return lpr;
}

bool aligned16(const void *p) { return ((uintptr_t)p & 15) == 0; }

template <typename S, bool CV, int LPR, int CPL>
hipError_t go_b1v(const BsrDev &A, const void *X, void *Y, const EpiArgs &e, int grid, hipStream_t st) {
    hipLaunchKernelGGL((k_spmm_b1v<S, CV, LPR, CPL>), dim3(grid), dim3(64 * kWaves), 0, st, A.nbr,
                       g_uniform_rows ? A.ulen : 0, A.row_ptr, A.col_ind, static_cast<const S *>(A.vals),
                       static_cast<const S *>(X), static_cast<S *>(Y), e);
    return hipGetLastError();
}

// The slab launch for lpr = b1v_lpr(...) (nonzero).
template <typename S, bool CV>
hipError_t by_b1v(int lpr, const BsrDev &A, const void *X, void *Y, int R, const EpiArgs &e, int grid,
                  hipStream_t st) {
    constexpr int VT = (int)sizeof(S) * (CV ? 2 : 1);
    auto pick = [&](auto bytes_c) -> hipError_t {
        constexpr int CPL = decltype(bytes_c)::value / VT;
        if constexpr (CPL < 1 || 16 % CPL != 0 || (CPL * VT) % 16 != 0) {
            return hipErrorInvalidValue;
        } else {
            if (R == 16) return go_b1v<S, CV, 16 / CPL, CPL>(A, X, Y, e, grid, st);
            if constexpr (32 / CPL <= 64) return go_b1v<S, CV, 32 / CPL, CPL>(A, X, Y, e, grid, st);
            return hipErrorInvalidValue;
        }
    };
    const int bytes = R / lpr * VT;
    if (bytes == 16) return pick(std::integral_constant<int, 16>{});
    if (bytes == 32) return pick(std::integral_constant<int, 32>{});
    if (bytes == 64) return pick(std::integral_constant<int, 64>{});
    return hipErrorInvalidValue;
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
// N is relabelled (col_of below) so that X, Y and the recurrence operands
// move as coalesced 16-byte loads/stores, and the (re, im) halves of a
// complex column sit in the same lane (tiles 2u, 2u + 1), so the complex
// dots need no shuffles.
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

// Column mapping of the MFMA SpMM: tile t of lane i holds real column
//   col(t, i) = (t / E) 16 E + i E + t % E,   E = min(NT, 16 / sizeof(S)),
// so every 16-byte load (E elements) of a row's 16 lanes covers 16 E
// contiguous columns, and the (re, im) halves of a complex column (E is even)
// stay in one lane, next to each other in t.
template <typename S, int NT> __host__ __device__ constexpr int col_epl() {
    return NT < 16 / (int)sizeof(S) ? NT : 16 / (int)sizeof(S);
}
template <typename S, int NT> __device__ __forceinline__ int col_of(int t, int i) {
    constexpr int E = col_epl<S, NT>();
    return (t / E) * 16 * E + i * E + t % E;
}

// 16-byte vector of the non-temporal builtins (which take native vector types)
typedef float f4nt __attribute__((ext_vector_type(4)));

// v[t] = row[col_of(t, i)]; NT: a non-temporal load (data read once, kept out
// of the way of the gathered rows in L2)
template <typename S, int NT, bool NTL = false>
__device__ __forceinline__ void load_cols(const S *__restrict__ row, int i, S (&v)[NT]) {
    constexpr int E = col_epl<S, NT>();
    constexpr int B = E * (int)sizeof(S);
#pragma unroll
    for (int q = 0; q < NT / E; ++q) {
        const S *p = row + q * 16 * E + i * E;
        if constexpr (B == 16) {
            if constexpr (NTL) {
                const f4nt w = __builtin_nontemporal_load(reinterpret_cast<const f4nt *>(p));
                __builtin_memcpy(&v[q * E], &w, 16);
            } else {
                const float4 w = *reinterpret_cast<const float4 *>(p);
                __builtin_memcpy(&v[q * E], &w, 16);
            }
        } else if constexpr (B == 8) {
            const float2 w = *reinterpret_cast<const float2 *>(p);
            __builtin_memcpy(&v[q * E], &w, 8);
        } else {
#pragma unroll
            for (int e = 0; e < E; ++e) v[q * E + e] = p[e];
        }
    }
}

// row[col_of(t, i)] = v[t], non-temporal for 16-byte pieces (an output row
// is not read again in the launch)
template <typename S, int NT>
__device__ __forceinline__ void store_cols(S *__restrict__ row, int i, const S (&v)[NT]) {
    constexpr int E = col_epl<S, NT>();
    constexpr int B = E * (int)sizeof(S);
#pragma unroll
    for (int q = 0; q < NT / E; ++q) {
        S *p = row + q * 16 * E + i * E;
        if constexpr (B == 16) {
            f4nt w;
            __builtin_memcpy(&w, &v[q * E], 16);
            __builtin_nontemporal_store(w, reinterpret_cast<f4nt *>(p));
        } else if constexpr (B == 8) {
            float2 w;
            __builtin_memcpy(&w, &v[q * E], 8);
            *reinterpret_cast<float2 *>(p) = w;
        } else {
#pragma unroll
            for (int e = 0; e < E; ++e) p[e] = v[q * E + e];
        }
    }
}

// Per-lane FP64 dot sums of the MFMA SpMM, ND dots x NT tiles. The
// precision form (S = float) keeps a compensation term beside every sum:
// its terms are products of two FP32 values, exact in FP64, so summing them
// with TwoSum is the compensated dot product Dot2 (Ogita, Rump and Oishi
// 2005) with an exact product step. The compensation terms live in LDS (a
// slot per lane, dot and tile; cl points at this lane's first), not in
// registers, so the kernel keeps its occupancy. The FP64 form sums with
// FMAs, as the reference's own dots do.
template <typename S, int ND, int NT> struct DotAcc {
    static constexpr bool kComp = sizeof(S) == 4;
    double s[ND][NT];
    double *cl = nullptr;
    __device__ __forceinline__ double &comp(int d, int t) { return cl[(d * NT + t) * 64]; }
    // comp(d, t) += x without reading it back (each lane owns its slots)
    __device__ __forceinline__ void comp_add(int d, int t, double x) {
        __hip_atomic_fetch_add(&cl[(d * NT + t) * 64], x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
    }
    __device__ __forceinline__ void zero() {
#pragma unroll
        for (int d = 0; d < ND; ++d)
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                s[d][t] = 0.0;
                if constexpr (kComp) {
                    if (cl) comp(d, t) = 0.0;
                }
            }
    }
    // s[d][t] += a1 b1 + a2 b2, in that order
    __device__ __forceinline__ void add2(int d, int t, double a1, double b1, double a2, double b2) {
        if constexpr (kComp) {
            double err = 0.0;
            two_sum_add(s[d][t], err, a1 * b1);
            two_sum_add(s[d][t], err, a2 * b2);
            comp_add(d, t, err);
        } else {
            s[d][t] = __builtin_fma(a2, b2, __builtin_fma(a1, b1, s[d][t]));
        }
    }
    __device__ __forceinline__ void add(int d, int t, double a, double b) {
        if constexpr (kComp) {
            double err = 0.0;
            two_sum_add(s[d][t], err, a * b);
            comp_add(d, t, err);
        } else {
            s[d][t] = __builtin_fma(a, b, s[d][t]);
        }
    }
    // the lane groups (lanes i, i + 16, i + 32, i + 48) of every tile folded
    // into one sum, its compensation included
    __device__ __forceinline__ double fold4(int d, int t) {
        double v = s[d][t];
        if constexpr (kComp) {
            double c = comp(d, t);
#pragma unroll
            for (int off = 16; off < 64; off <<= 1) {
                const double o = __shfl_xor(v, off), oc = __shfl_xor(c, off);
                two_sum_add(v, c, o);
                c += oc;
            }
            return v + c;
        } else {
            v += __shfl_xor(v, 16);
            v += __shfl_xor(v, 32);
            return v;
        }
    }
};

template <typename S, int NT, int MODE>
using MfmaDots = DotAcc<S, (epi_ndot(MODE) > 0 ? epi_ndot(MODE) : 1), NT>;

// With PRE (the FP64 step), pxs and pxp arrive holding the Xs and Xp values
// of the first output rows (loaded by the kernel before the row's K loop), and
// each pass loads the next pass's values into them as soon as its own are
// consumed, so the epilogue's reads overlap its compute instead of stalling.
template <typename S, int NT, int MODE, bool CV, bool PRE = false>
__device__ __forceinline__ void mfma_epilogue(const typename Acc4<S>::T (&acc)[NT], int64_t br,
                                              int b, int W, int col0, int i, int g,
                                              const EpiArgs &e, MfmaDots<S, NT, MODE> &dacc,
                                              S *__restrict__ Y, S (&pxs)[NT], S (&pxp)[NT]) {
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
        const int64_t base = (br * b + r) * (int64_t)W + col0;
        S y[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            // element v of acc[t], selected without dynamic indexing
            y[t] = v == 0 ? acc[t][0] : v == 1 ? acc[t][1] : v == 2 ? acc[t][2] : acc[t][3];
        }
        // the next pass's output row of this lane (PRE loads it early)
        const int rn = sizeof(S) == 8 ? 4 * (v + 1) + g : 4 * g + v + 1;
        const bool next = PRE && v < 3 && rn < b;
        const int64_t base_n = (br * b + rn) * (int64_t)W + col0;
        if constexpr (MODE != EPI_PLAIN) {
            S xs[NT];
            if constexpr (PRE) {
#pragma unroll
                for (int t = 0; t < NT; ++t) xs[t] = pxs[t];
            } else {
                load_cols<S, NT, true>(XS + base, i, xs);
            }
#pragma unroll
            for (int t = 0; t < NT; ++t)
                y[t] = (S)__builtin_fma(e.alpha, (double)y[t], e.beta * (double)xs[t]);
            if constexpr (MODE == EPI_STEP || MODE == EPI_PLAINSTEP || MODE == EPI_ADJ) {
                if constexpr (PRE) {
#pragma unroll
                    for (int t = 0; t < NT; ++t) y[t] = y[t] - pxp[t];
                    if (next) load_cols<S, NT, true>(XP + base_n, i, pxp);
                } else {
                    S xp[NT];
                    load_cols<S, NT, true>(XP + base, i, xp);
#pragma unroll
                    for (int t = 0; t < NT; ++t) y[t] = y[t] - xp[t];
                }
            }
            if constexpr (MODE == EPI_ADJ) {
                const S *qs[3] = {static_cast<const S *>(e.qm), static_cast<const S *>(e.qz),
                                  static_cast<const S *>(e.qp)};
                const double cs[3] = {e.cm, e.cz, e.cp};
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    S qv[NT];
                    load_cols<S, NT, true>(qs[q] + base, i, qv);
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
                        dacc.add2(d, t, ai, bi, ar, br_);
                        if (!self) dacc.add2(d, t + 1, -ai, br_, ar, bi);
                    }
                } else {
#pragma unroll
                    for (int t = 0; t < NT; ++t) dacc.add(d, t, (double)a[t], (double)bb[t]);
                }
            };
            if constexpr (MODE == EPI_FIRST) {
                dot(0, xs, xs, true);
                dot(1, xs, y, false);
                dot(2, y, y, true);
            } else if constexpr (MODE == EPI_STEP) {
                dot(0, y, y, true);
                dot(1, y, xs, false);
                if constexpr (PRE) {
                    if (next) load_cols<S, NT, true>(XS + base_n, i, pxs);
                }
            } else if constexpr (MODE == EPI_PLAINSTEP) {
                S ps[NT];
                load_cols<S, NT, true>(PSI + base, i, ps);
                dot(0, ps, y, false);
            }
        }
        store_cols<S, NT>(Y + base, i, y);
    }
}

// Per-workgroup dot partials of the MFMA SpMMs: the lane groups of every
// column are folded, then the waves sharing a column chunk, into
// e.partials[(blockIdx.x * NDOT + d) * R + u].
template <typename S, int NT, int MODE, bool CV>
__device__ __forceinline__ void mfma_dot_partials(MfmaDots<S, NT, MODE> &dacc, int wave, int i, int g,
                                                  int nchunks, int R, const EpiArgs &e) {
    constexpr int NDOT = epi_ndot(MODE);
    constexpr int CW = 16 * NT;
    if constexpr (NDOT > 0) {
        // per wave: CW real columns -> CW/2 complex (or CW real) units
        constexpr int UPC = CV ? CW / 2 : CW;
        __shared__ double2 sm[kWaves][UPC];
        // unrolled: the sums stay in registers (a loop over d would index them)
#pragma unroll
        for (int d = 0; d < NDOT; ++d) {
            // fold the four lane groups (rows of the 16-row tile) of every column
            double f[NT];
#pragma unroll
            for (int t = 0; t < NT; ++t) f[t] = dacc.fold4(d, t);
            if (g == 0) {
#pragma unroll
                for (int t = 0; t < NT; t += (CV ? 2 : 1)) {
                    const int uu = CV ? col_of<S, NT>(t, i) / 2 : col_of<S, NT>(t, i);
                    sm[wave][uu] = CV ? double2{f[t], f[t + 1]} : double2{f[t], 0.0};
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
                                                   int W, int R, int xcd_map,
                                                   const int32_t *__restrict__ sched, int sreg,
                                                   EpiArgs e) {
    using A4 = typename Acc4<S>::T;
    constexpr int NDOT = MODE == EPI_FIRST ? 3 : MODE == EPI_STEP ? 2 : MODE == EPI_PLAINSTEP ? 1 : 0;
    constexpr int CW = 16 * NT; // real columns per chunk
    // the FP64 step at sixteen tiles (one wave per SIMD, registers to spare)
    // loads its epilogue's inputs ahead
    constexpr bool PRE = sizeof(S) == 8 && NT == 16 && MODE == EPI_STEP;

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
    // With a row schedule the ranges are whole regions of sreg positions
    // (a region is what one XCD keeps in flight), and position p computes
    // block row sched[p].
    int64_t r_lo = 0, r_hi = nbr;
    int64_t kq = blockIdx.x, nk = gridDim.x;
    if (xcd_map) {
        const int x = blockIdx.x & 7;
        kq = blockIdx.x >> 3;
        nk = gridDim.x >> 3;
        if (sched) {
            const int64_t nreg = (nbr + sreg - 1) / sreg;
            r_lo = min(nbr, nreg * x / 8 * sreg);
            r_hi = min(nbr, nreg * (x + 1) / 8 * sreg);
        } else {
            r_lo = nbr * x / 8;
            r_hi = nbr * (x + 1) / 8;
        }
    }
    const int64_t total = (r_hi - r_lo) * nchunks;

    MfmaDots<S, NT, MODE> dacc;
    if constexpr (MfmaDots<S, NT, MODE>::kComp && NDOT > 0) {
        __shared__ double comp_lds[kWaves * NDOT * NT * 64];
        dacc.cl = comp_lds + wave * NDOT * NT * 64 + lane;
    }
    dacc.zero();

    for (int64_t it = kq * kWaves + wave; it < total; it += nk * kWaves) {
        const int64_t pos = r_lo + it / nchunks;
        const int64_t br = sched ? (int64_t)sched[pos] : pos;
        const int64_t s0 = row_ptr[br];
        const int nblk = (int)(row_ptr[br + 1] - s0);
        const int nsteps = (nblk * b + 3) >> 2;
        // the first output rows' Xs, Xp values, in flight during the K loop
        S pxs[NT], pxp[NT];
        if constexpr (PRE) {
            const int64_t base0 = (br * b + g) * (int64_t)W + col0; // FP64 pass 0: row g < 8 <= b
            load_cols<S, NT, true>(static_cast<const S *>(e.xs) + base0, i, pxs);
            load_cols<S, NT, true>(static_cast<const S *>(e.xp) + base0, i, pxp);
        }

        A4 acc[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[t] = A4{0, 0, 0, 0};

        // Step st covers K columns 4 st .. 4 st + 3 of the row's column-major
        // b x K panel (element (i, k) at s0 b^2 + k b + i); lane group g takes
        // column k = 4 st + g, its X row from the gather table (wave-uniform
        // scalar loads, so no vector load sits between the steps' loads). PD
        // register buffers: steps st + 1 .. st + PD - 1 are in flight while
        // step st's MFMAs run. Past the row's K the table still names a valid row
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
            load_cols<S, NT>(X + xr * W + col0, i, bv);
            // the next step's table entries (in range: the table keeps 8 PD - 2 <= 64 spare entries)
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

        mfma_epilogue<S, NT, MODE, CV, PRE>(acc, br, b, W, col0, i, g, e, dacc, Y, pxs, pxp);
    }

    mfma_dot_partials<S, NT, MODE, CV>(dacc, wave, i, g, nchunks, R, e);
}

// Workgroups of one MFMA SpMM kernel that are resident at once, at most two
// per compute unit, a multiple of 8 (whole XCD rounds), at most kMaxGrid.
// With no queued workgroups the grid-stride loop walks the schedule's regions
// in order on every XCD.
int mfma_resident_grid(const void *kernel) {
    int dev = 0, ncu = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu < 1 ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 64 * kWaves, 0) != hipSuccess || per_cu < 1)
        return 512;
    const int64_t g = ((int64_t)std::min(per_cu, 2) * ncu / 8) * 8;
    return (int)std::max<int64_t>(8, std::min<int64_t>(g, kMaxGrid));
}

// Set while spmm_launch_grid asks the MFMA dispatch which grid it would take:
// launch_mfma then reports the grid without launching.
thread_local bool t_grid_query = false;

int mfma_grid(int grid_req, int resident) {
    const int c = g_mfma_grid_cap > 0 ? g_mfma_grid_cap : resident;
    return grid_req < c ? grid_req : c;
}

template <typename S, int NT, int MODE, bool CV, int PD>
hipError_t launch_mfma(const BsrDev &A, const void *X, void *Y, int W, int R, const EpiArgs &e, int grid_req,
                       hipStream_t st, int *grid_out) {
    static std::atomic<int> resident{0};
    int res = resident.load(std::memory_order_relaxed);
    if (res == 0) {
        res = mfma_resident_grid(reinterpret_cast<const void *>(&k_spmm_mfma<S, NT, MODE, CV, PD>));
        resident.store(res, std::memory_order_relaxed);
    }
    const int grid = mfma_grid(grid_req, res);
    if (grid_out) *grid_out = grid;
    if (t_grid_query) return hipSuccess;
    hipLaunchKernelGGL((k_spmm_mfma<S, NT, MODE, CV, PD>), dim3(grid), dim3(64 * kWaves), 0, st, A.nbr,
                       (int)A.b, A.row_ptr, A.xrow, static_cast<const S *>(A.vals), static_cast<const S *>(X),
                       static_cast<S *>(Y), W, R, g_xcd_map, A.sched, A.sreg, e);
    return hipGetLastError();
}

template <typename S, int NT, bool CV, int PD>
hipError_t go_mfma_pd(const BsrDev &A, const void *X, void *Y, int W, int R, const EpiArgs &e,
                      int grid, hipStream_t st, int *grid_out) {
    switch (e.mode) {
    case EPI_PLAIN: return launch_mfma<S, NT, EPI_PLAIN, CV, PD>(A, X, Y, W, R, e, grid, st, grid_out);
    case EPI_FIRST: return launch_mfma<S, NT, EPI_FIRST, CV, PD>(A, X, Y, W, R, e, grid, st, grid_out);
    case EPI_STEP: return launch_mfma<S, NT, EPI_STEP, CV, PD>(A, X, Y, W, R, e, grid, st, grid_out);
    case EPI_PLAINSTEP: return launch_mfma<S, NT, EPI_PLAINSTEP, CV, PD>(A, X, Y, W, R, e, grid, st, grid_out);
    default: return launch_mfma<S, NT, EPI_ADJ, CV, PD>(A, X, Y, W, R, e, grid, st, grid_out);
    }
}

// Sixteen tiles per wave over complex vectors (R = 128 complex): the plain
// product or the step at prefetch depth pd (2 .. 6).
template <typename S, int MODE>
hipError_t go_mfma16(int pd, const BsrDev &A, const void *X, void *Y, int W, int R, const EpiArgs &e, int grid,
                     hipStream_t st, int *grid_out) {
    switch (pd) {
    case 2: return launch_mfma<S, 16, MODE, true, 2>(A, X, Y, W, R, e, grid, st, grid_out);
    case 4: return launch_mfma<S, 16, MODE, true, 4>(A, X, Y, W, R, e, grid, st, grid_out);
    case 5: return launch_mfma<S, 16, MODE, true, 5>(A, X, Y, W, R, e, grid, st, grid_out);
    case 6: return launch_mfma<S, 16, MODE, true, 6>(A, X, Y, W, R, e, grid, st, grid_out);
    default: return launch_mfma<S, 16, MODE, true, 3>(A, X, Y, W, R, e, grid, st, grid_out);
    }
}

// register prefetch depth of k_spmm_mfma (g_mfma_pd64, g_mfma_pd32, g_pd16)
template <typename S, int NT, bool CV>
hipError_t go_mfma(const BsrDev &A, const void *X, void *Y, int W, int R, const EpiArgs &e,
                   int grid, hipStream_t st, int *grid_out) {
    if constexpr (NT == 16 && CV) {
        const int ty = sizeof(S) == 8 ? 0 : 1;
        if (e.mode == EPI_PLAIN)
            return go_mfma16<S, EPI_PLAIN>(g_pd16[ty][0], A, X, Y, W, R, e, grid, st, grid_out);
        if (e.mode == EPI_STEP)
            return go_mfma16<S, EPI_STEP>(g_pd16[ty][1], A, X, Y, W, R, e, grid, st, grid_out);
    }
    if constexpr (sizeof(S) == 8) {
        if (g_mfma_pd64 == 3 && e.mode == EPI_PLAIN) return go_mfma_pd<S, NT, CV, 3>(A, X, Y, W, R, e, grid, st, grid_out);
        return go_mfma_pd<S, NT, CV, 2>(A, X, Y, W, R, e, grid, st, grid_out);
    } else {
        if (g_mfma_pd32 == 4) return go_mfma_pd<S, NT, CV, 4>(A, X, Y, W, R, e, grid, st, grid_out);
        return go_mfma_pd<S, NT, CV, 3>(A, X, Y, W, R, e, grid, st, grid_out);
    }
}

template <typename S, bool CV>
hipError_t by_nt(int nt, const BsrDev &A, const void *X, void *Y, int W, int R, const EpiArgs &e,
                 int grid, hipStream_t st, int *grid_out) {
    switch (nt) {
    case 2: return go_mfma<S, 2, CV>(A, X, Y, W, R, e, grid, st, grid_out);
    case 4: return go_mfma<S, 4, CV>(A, X, Y, W, R, e, grid, st, grid_out);
    case 8: return go_mfma<S, 8, CV>(A, X, Y, W, R, e, grid, st, grid_out);
    default: return go_mfma<S, 16, CV>(A, X, Y, W, R, e, grid, st, grid_out);
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

// One (dot, column) pair per wave; blockIdx.y selects a step of a batched
// reduction (partials and out advance by their step strides).
__global__ __launch_bounds__(256) void k_reduce(const double2 *__restrict__ partials, int ngroups,
                                                int ndot, int R, double2 *__restrict__ out, int64_t pstride,
                                                int64_t ostride) {
    partials += (int64_t)blockIdx.y * pstride;
    out += (int64_t)blockIdx.y * ostride;
    const int lane = threadIdx.x & 63;
    const int pair = blockIdx.x * (int)(blockDim.x >> 6) + (threadIdx.x >> 6);
    const int npairs = ndot * R;
    if (pair >= npairs) return;
    const int d = pair / R;
    const int u = pair - d * R;
    // compensated FP64 sums over the workgroup partials, then a fixed xor tree
    double sx = 0.0, cx = 0.0, sy = 0.0, cy = 0.0;
    // 32 loads in flight per lane (all of a 2048-workgroup launch's partials
    // in one round trip) before they are summed in order
    constexpr int B = 32;
    for (int g0 = lane; g0 < ngroups; g0 += 64 * B) {
        double2 p[B];
#pragma unroll
        for (int q = 0; q < B; ++q) {
            const int gi = g0 + 64 * q;
            p[q] = gi < ngroups ? partials[((int64_t)gi * ndot + d) * R + u] : double2{0.0, 0.0};
        }
#pragma unroll
        for (int q = 0; q < B; ++q) {
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

// Block gather: one block of m scalars per row of threads' grid-stride walk.
__global__ __launch_bounds__(256) void k_block_perm(const double *__restrict__ in,
                                                    const int32_t *__restrict__ perm,
                                                    double *__restrict__ out, int64_t nblk, int64_t m) {
    const int64_t n = nblk * m;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t p = t / m;
        out[t] = in[(int64_t)perm[p] * m + (t - p * m)];
    }
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

namespace {
// Workgroups of the VALU SpMM: kWaves waves of 64 / lpr block rows each.
int valu_grid(int64_t nbr, int lpr) {
    const int G = 64 / lpr;
    const int64_t rows_per_wg = (int64_t)kWaves * G;
    int64_t g = (nbr + rows_per_wg - 1) / rows_per_wg;
    if (g > kMaxGrid) g = kMaxGrid;
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

int spmm_launch_grid(const BsrDev &A, VecMode vm, int R) {
    const int W = R * (vm.complex_vec ? 2 : 1);
    int nt = mfma_nt(W);
    while (nt > 2 && nt > g_mfma_ntcap) nt /= 2;
    const int nchunks = nt > 0 ? W / (16 * nt) : 0;
    if (!A.is_complex && A.b >= 8 && A.xrow && nt > 0 && kWaves % nchunks == 0) {
        // the grid the step launch of this operand takes (kernels differ in
        // their resident capacity)
        EpiArgs e;
        e.mode = EPI_STEP;
        int g = 0;
        t_grid_query = true;
        const hipError_t he = launch_spmm(A, vm, nullptr, nullptr, R, e, nullptr, &g);
        t_grid_query = false;
        return he == hipSuccess ? g : kMaxGrid;
    }
    return valu_grid(A.nbr, valu_lpr(A, vm.fp32, vm.complex_vec, R));
}

hipError_t launch_spmm(const BsrDev &A, VecMode vm, const void *X, void *Y, int R,
                       const EpiArgs &e, hipStream_t stream, int *grid_out) {
    if (R < 1 || R > 256 || A.b < 1 || A.b > 16) return hipErrorInvalidValue;
    if (A.is_complex && !vm.complex_vec) return hipErrorInvalidValue;
    const int W = R * (vm.complex_vec ? 2 : 1);
    int nt = mfma_nt(W);
    while (nt > 2 && nt > g_mfma_ntcap) nt /= 2;
    const int nchunks = nt > 0 ? W / (16 * nt) : 0;
    if (!A.is_complex && A.b >= 8 && A.xrow && nt > 0 && kWaves % nchunks == 0) {
        int64_t g = (A.nbr * nchunks + kWaves - 1) / kWaves;
        if (g > kMaxGrid) g = kMaxGrid;
        g = ((g + 7) / 8) * 8; // whole XCD rounds (kMaxGrid is a multiple of 8)
        const int grid = (int)g; // capped further by mfma_grid
        if (vm.fp32) {
            return vm.complex_vec ? by_nt<float, true>(nt, A, X, Y, W, R, e, grid, stream, grid_out)
                                  : by_nt<float, false>(nt, A, X, Y, W, R, e, grid, stream, grid_out);
        }
        return vm.complex_vec ? by_nt<double, true>(nt, A, X, Y, W, R, e, grid, stream, grid_out)
                              : by_nt<double, false>(nt, A, X, Y, W, R, e, grid, stream, grid_out);
    }
    // sized by the layout the launch takes (the wide b = 1 layout covers twice
    // the rows per workgroup)
    const int grid = valu_grid(A.nbr, valu_lpr(A, vm.fp32, vm.complex_vec, R));
    if (grid_out) *grid_out = grid;
    if (const int lpr = b1v_lpr(A, vm.fp32, vm.complex_vec, R)) {
        const void *ops[] = {X, Y, e.xs, e.xp, e.psi, e.qm, e.qz, e.qp};
        bool ok = true;
        for (const void *o : ops) ok = ok && aligned16(o);
        if (ok) {
            if (vm.fp32) return vm.complex_vec ? by_b1v<float, true>(lpr, A, X, Y, R, e, grid, stream)
                                               : by_b1v<float, false>(lpr, A, X, Y, R, e, grid, stream);
            return vm.complex_vec ? by_b1v<double, true>(lpr, A, X, Y, R, e, grid, stream)
                                  : by_b1v<double, false>(lpr, A, X, Y, R, e, grid, stream);
        }
    }
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
void set_mfma_grid_cap(int g) { g_mfma_grid_cap = g >= 8 && g <= kMaxGrid ? (g / 8) * 8 : 0; }
void set_b1_wide(int mode) { g_b1_wide = mode == 0 || mode == 1 ? mode : 2; }
void set_uniform_rows(int on) { g_uniform_rows = on != 0; }
void set_b1_slab(int bytes) { g_b1v_bytes = bytes == 0 || bytes == 16 || bytes == 32 || bytes == 64 ? bytes : -1; }
void set_b1_slab_min_grid(int g) { g_b1v_min_grid = g >= 0 ? g : 512; }
void set_mfma_pd32(int p) { g_mfma_pd32 = p == 4 ? 4 : 3; }
void set_mfma_pd16(int fp32, int step, int p) {
    if (p >= 2 && p <= 6) g_pd16[fp32 ? 1 : 0][step ? 1 : 0] = p;
}
void set_mfma_pd64(int p) { g_mfma_pd64 = p == 2 ? 2 : 3; }

hipError_t launch_block_transpose(const double *in, double *out, int64_t nnzb, int b, bool cplx,
                                  hipStream_t stream) {
    hipLaunchKernelGGL(k_block_transpose, dim3(grid_1d(nnzb * b * b)), dim3(256), 0, stream, in,
                       out, nnzb, b, cplx ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_reduce(const double2 *partials, int ngroups, int ndot, int R, double2 *out,
                         hipStream_t stream) {
    return launch_reduce_steps(partials, 0, ngroups, ndot, R, out, 0, 1, stream);
}

hipError_t launch_reduce_steps(const double2 *partials, int64_t pstride, int ngroups, int ndot, int R,
                               double2 *out, int64_t ostride, int nsteps, hipStream_t stream) {
    if (nsteps < 1) return hipSuccess;
    // one wave (one dot and column) a workgroup: the strided partial loads
    // spread over as many compute units as there are sums
    const int npairs = ndot * R;
    hipLaunchKernelGGL(k_reduce, dim3(npairs, (unsigned)nsteps), dim3(64), 0, stream, partials, ngroups, ndot, R,
                       out, pstride, ostride);
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

hipError_t launch_block_perm(const double *in, const int32_t *perm, double *out, int64_t nblk, int64_t m,
                             hipStream_t stream) {
    if (nblk <= 0 || m <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_block_perm, dim3(grid_1d(nblk * m)), dim3(256), 0, stream, in, perm, out, nblk, m);
    return hipGetLastError();
}

hipError_t launch_stage(const double *in, bool in_complex, void *out, bool out_complex,
                        bool out_fp32, int64_t n, hipStream_t stream) {
    hipLaunchKernelGGL(k_stage, dim3(grid_1d(n)), dim3(256), 0, stream, in, in_complex ? 1 : 0,
                       out, out_complex ? 1 : 0, out_fp32 ? 1 : 0, n);
    return hipGetLastError();
}

} // namespace rg
