// Graph-local S^{-1/2} (the seed of rsatb's inverse_sqrt,
// rsatb/calc/isqrtm.hpp:137-232) as batched dense FP64 Newton-Schulz.
//
// For block row i the patch is A_i = S[C_i, C_i], C_i = the stored block
// columns of row i (ascending), n_i = b |C_i|; the factor's block row i is
// the rows of A_i^{-1/2} that belong to block i (isqrtm.hpp:140-146: only one
// block row of every subgraph's inverse square root is used).
//
// The coupled Newton-Schulz iteration (Higham, Functions of Matrices, 2008,
// eq. 6.35) on Y0 = A/c, Z0 = I,
//     T_k = (3 I - Z_k Y_k) / 2,  Y_{k+1} = Y_k T_k,  Z_{k+1} = T_k Z_k,
// converges quadratically to Y -> (A/c)^{1/2}, Z -> (A/c)^{-1/2} when the
// spectrum of A/c lies in (0, 1]; c is the Gershgorin bound of the patch
// (max absolute row sum >= lambda_max). Patches are padded to ld x ld
// (ld a multiple of 64) with the identity, which the iteration keeps fixed.
// Every GEMM is a 64 x 64 output tile per workgroup on v_mfma_f64_16x16x4f64
// with 64 x 16 / 16 x 64 operand panels staged in LDS.
#include "rg_internal.h"

namespace rg {
namespace {

typedef double D4 __attribute__((ext_vector_type(4)));

// block (r, c) of a BSR with sorted columns: its index, or -1
__device__ __forceinline__ int64_t find_block(const int64_t *__restrict__ row_ptr,
                                              const int32_t *__restrict__ col_ind, int64_t r,
                                              int32_t c) {
    int64_t lo = row_ptr[r], hi = row_ptr[r + 1] - 1;
    while (lo <= hi) {
        const int64_t mid = (lo + hi) >> 1;
        const int32_t v = col_ind[mid];
        if (v == c) return mid;
        if (v < c) lo = mid + 1;
        else hi = mid - 1;
    }
    return -1;
}

// One workgroup per patch: Y = diag(A/c, I), Z = I; cs[p] = c.
__global__ __launch_bounds__(256) void k_patch_gather(int64_t first, int b,
                                                      const int64_t *__restrict__ row_ptr,
                                                      const int32_t *__restrict__ col_ind,
                                                      const double *__restrict__ vals, int ld,
                                                      double *__restrict__ Y,
                                                      double *__restrict__ Z,
                                                      double *__restrict__ cs) {
    const int64_t p = blockIdx.x;
    const int64_t row = first + p;
    const int64_t s0 = row_ptr[row];
    const int nb = (int)(row_ptr[row + 1] - s0);
    const int n = nb * b;
    double *__restrict__ Yp = Y + p * (int64_t)ld * ld;
    double *__restrict__ Zp = Z + p * (int64_t)ld * ld;
    const int64_t ldd = (int64_t)ld * ld;
    for (int64_t t = threadIdx.x; t < ldd; t += blockDim.x) {
        const int r = (int)(t / ld), c = (int)(t - (t / ld) * ld);
        const double d = (r == c) ? 1.0 : 0.0;
        Zp[t] = d;
        Yp[t] = (r >= n) ? d : 0.0; // padding rows keep the identity
    }
    __syncthreads();
    // Gershgorin bound: max_r sum_c |A_rc|
    __shared__ double red[256];
    double mymax = 0.0;
    for (int r = threadIdx.x; r < n; r += blockDim.x) {
        const int pb = r / b, rr = r - pb * b;
        const int32_t ci = col_ind[s0 + pb];
        double sum = 0.0;
        for (int q = 0; q < nb; ++q) {
            const int64_t sb = find_block(row_ptr, col_ind, ci, col_ind[s0 + q]);
            if (sb < 0) continue;
            const double *blk = vals + sb * (int64_t)b * b; // column-major
            for (int cc = 0; cc < b; ++cc) sum += fabs(blk[cc * b + rr]);
        }
        mymax = fmax(mymax, sum);
    }
    red[threadIdx.x] = mymax;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + off]);
        __syncthreads();
    }
    const double c = red[0] > 0.0 ? red[0] : 1.0;
    if (threadIdx.x == 0) cs[p] = c;
    const double inv = 1.0 / c;
    for (int r = threadIdx.x; r < n; r += blockDim.x) {
        const int pb = r / b, rr = r - pb * b;
        const int32_t ci = col_ind[s0 + pb];
        for (int q = 0; q < nb; ++q) {
            const int64_t sb = find_block(row_ptr, col_ind, ci, col_ind[s0 + q]);
            if (sb < 0) continue;
            const double *blk = vals + sb * (int64_t)b * b;
            for (int cc = 0; cc < b; ++cc) Yp[(int64_t)r * ld + q * b + cc] = blk[cc * b + rr] * inv;
        }
    }
}

// C = A B per patch (MODE 1), or C = 1.5 I - 0.5 A B with
// resid[p] += sum (I - A B)^2 (MODE 0). ld x ld row-major, ld % 64 == 0.
template <int MODE>
__global__ __launch_bounds__(256) void k_ns_gemm(int ld, const double *__restrict__ A,
                                                 const double *__restrict__ B,
                                                 double *__restrict__ C,
                                                 double *__restrict__ resid) {
    const int tiles = ld / 64;
    const int64_t p = blockIdx.x / (tiles * tiles);
    const int tile = (int)(blockIdx.x - p * tiles * tiles);
    const int row0 = (tile / tiles) * 64, col0 = (tile % tiles) * 64;
    const int64_t off = p * (int64_t)ld * ld;
    const double *__restrict__ Ap = A + off;
    const double *__restrict__ Bp = B + off;
    double *__restrict__ Cp = C + off;

    __shared__ double As[64][17];
    __shared__ double Bs[16][64];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int i = lane & 15, g = lane >> 4;
    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
    D4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int c = 0; c < 2; ++c) acc[a][c] = D4{0, 0, 0, 0};

    for (int kb = 0; kb < ld; kb += 16) {
        {
            const int r = tid >> 2, c4 = (tid & 3) * 4;
            const double *src = Ap + (int64_t)(row0 + r) * ld + kb + c4;
#pragma unroll
            for (int q = 0; q < 4; ++q) As[r][c4 + q] = src[q];
            const int br = tid >> 4, bc = (tid & 15) * 4;
            const double *srcb = Bp + (int64_t)(kb + br) * ld + col0 + bc;
#pragma unroll
            for (int q = 0; q < 4; ++q) Bs[br][bc + q] = srcb[q];
        }
        __syncthreads();
#pragma unroll
        for (int k4 = 0; k4 < 16; k4 += 4) {
            double av[2], bv[2];
#pragma unroll
            for (int a = 0; a < 2; ++a) av[a] = As[wr + a * 16 + i][k4 + g];
#pragma unroll
            for (int c = 0; c < 2; ++c) bv[c] = Bs[k4 + g][wc + c * 16 + i];
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int c = 0; c < 2; ++c)
                    acc[a][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a], bv[c], acc[a][c], 0, 0, 0);
        }
        __syncthreads();
    }

    double rs = 0.0;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r = row0 + wr + a * 16 + 4 * v + g; // f64 accumulator layout
                const int cc = col0 + wc + c * 16 + i;
                const double x = acc[a][c][v];
                if (MODE == 0) {
                    const double d = (r == cc ? 1.0 : 0.0) - x;
                    rs = __builtin_fma(d, d, rs);
                    Cp[(int64_t)r * ld + cc] = (r == cc ? 1.5 : 0.0) - 0.5 * x;
                } else {
                    Cp[(int64_t)r * ld + cc] = x;
                }
            }
    if (MODE == 0) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) rs += __shfl_xor(rs, o);
        if (lane == 0) atomicAdd(resid + p, rs);
    }
}

// Z block row i = rows of block i in Zp / sqrt(c), into column-major blocks.
__global__ __launch_bounds__(256) void k_patch_extract(int64_t first, int b,
                                                       const int64_t *__restrict__ row_ptr,
                                                       const int32_t *__restrict__ col_ind,
                                                       int ld, const double *__restrict__ Z,
                                                       const double *__restrict__ cs,
                                                       double *__restrict__ out) {
    const int64_t p = blockIdx.x;
    const int64_t row = first + p;
    const int64_t s0 = row_ptr[row];
    const int nb = (int)(row_ptr[row + 1] - s0);
    // position of block i within its own pattern
    int me = -1;
    for (int q = 0; q < nb; ++q)
        if (col_ind[s0 + q] == (int32_t)row) me = q;
    if (me < 0) return; // rows without a diagonal block are rejected on the host
    const double sc = 1.0 / sqrt(cs[p]);
    const double *__restrict__ Zp = Z + p * (int64_t)ld * ld;
    const int per = b * b;
    for (int t = threadIdx.x; t < nb * per; t += blockDim.x) {
        const int q = t / per, e = t - q * per;
        const int cc = e / b, rr = e - cc * b; // column-major element (rr, cc)
        out[(s0 + q) * (int64_t)per + e] = Zp[(int64_t)(me * b + rr) * ld + q * b + cc] * sc;
    }
}

// out block (i, j) = conj(in block (j, i))^T, for a pattern that is
// symmetric; column-major blocks, real or complex.
__global__ __launch_bounds__(256) void k_bsr_ct_vals(int64_t nbr, int b, int cplx,
                                                     const int64_t *__restrict__ row_ptr,
                                                     const int32_t *__restrict__ col_ind,
                                                     const double *__restrict__ in,
                                                     double *__restrict__ out) {
    const int per = b * b;
    const int64_t nnzb = row_ptr[nbr];
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nnzb * per;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t s = t / per;
        const int e = (int)(t - s * per);
        const int cc = e / b, rr = e - cc * b;
        // row of block s: binary search in row_ptr
        int64_t lo = 0, hi = nbr - 1;
        while (lo < hi) {
            const int64_t mid = (lo + hi + 1) >> 1;
            if (row_ptr[mid] <= s) lo = mid;
            else hi = mid - 1;
        }
        const int64_t i = lo;
        const int32_t j = col_ind[s];
        const int64_t s2 = find_block(row_ptr, col_ind, j, (int32_t)i);
        // out(rr, cc) = conj(in_{ji}(cc, rr)); in is column-major: (cc, rr) at rr*b + cc
        const int64_t src = s2 * per + (int64_t)rr * b + cc;
        if (cplx) {
            out[2 * t] = in[2 * src];
            out[2 * t + 1] = -in[2 * src + 1];
        } else {
            out[t] = in[src];
        }
    }
}

} // namespace

hipError_t launch_patch_gather(int64_t first, int64_t count, int b, const int64_t *row_ptr,
                               const int32_t *col_ind, const double *vals, int ld, double *Y,
                               double *Z, double *cs, hipStream_t st) {
    hipLaunchKernelGGL(k_patch_gather, dim3((unsigned)count), dim3(256), 0, st, first, b, row_ptr,
                       col_ind, vals, ld, Y, Z, cs);
    return hipGetLastError();
}

hipError_t launch_ns_gemm(int mode, int64_t count, int ld, const double *A, const double *B,
                          double *C, double *resid, hipStream_t st) {
    const int tiles = ld / 64;
    const dim3 grid((unsigned)(count * tiles * tiles));
    if (mode == 0)
        hipLaunchKernelGGL(k_ns_gemm<0>, grid, dim3(256), 0, st, ld, A, B, C, resid);
    else
        hipLaunchKernelGGL(k_ns_gemm<1>, grid, dim3(256), 0, st, ld, A, B, C, resid);
    return hipGetLastError();
}

hipError_t launch_patch_extract(int64_t first, int64_t count, int b, const int64_t *row_ptr,
                                const int32_t *col_ind, int ld, const double *Z, const double *cs,
                                double *out, hipStream_t st) {
    hipLaunchKernelGGL(k_patch_extract, dim3((unsigned)count), dim3(256), 0, st, first, b, row_ptr,
                       col_ind, ld, Z, cs, out);
    return hipGetLastError();
}

hipError_t launch_bsr_ct_vals(int64_t nbr, int b, bool cplx, int64_t nnzb, const int64_t *row_ptr,
                              const int32_t *col_ind, const double *in, double *out,
                              hipStream_t st) {
    int64_t g = (nnzb * b * b + 255) / 256;
    if (g > 16384) g = 16384;
    if (g < 1) g = 1;
    hipLaunchKernelGGL(k_bsr_ct_vals, dim3((unsigned)g), dim3(256), 0, st, nbr, b, cplx ? 1 : 0,
                       row_ptr, col_ind, in, out);
    return hipGetLastError();
}

} // namespace rg
