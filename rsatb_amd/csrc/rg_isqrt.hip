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

// Per patch, ld x ld row-major, ld % 64 == 0:
//   MODE 0: C = 1.5 I - 0.5 A B, resid[p] += sum (I - A B)^2
//   MODE 1: C = A B
//   MODE 2: C = -0.5 (A B + A2 B2)
//   MODE 3: C = A B + A2 B2
// (modes 2 and 3 are the tangent products of the coupled iteration).
template <int MODE>
__global__ __launch_bounds__(256) void k_ns_gemm(int ld, const double *__restrict__ A,
                                                 const double *__restrict__ B,
                                                 const double *__restrict__ A2,
                                                 const double *__restrict__ B2,
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

    constexpr int NPROD = MODE >= 2 ? 2 : 1;
    for (int kq = 0; kq < NPROD * ld; kq += 16) {
        const int kb = kq < ld ? kq : kq - ld;
        const double *__restrict__ Aq = kq < ld ? Ap : A2 + off;
        const double *__restrict__ Bq = kq < ld ? Bp : B2 + off;
        {
            const int r = tid >> 2, c4 = (tid & 3) * 4;
            const double *src = Aq + (int64_t)(row0 + r) * ld + kb + c4;
#pragma unroll
            for (int q = 0; q < 4; ++q) As[r][c4 + q] = src[q];
            const int br = tid >> 4, bc = (tid & 15) * 4;
            const double *srcb = Bq + (int64_t)(kb + br) * ld + col0 + bc;
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
                } else if (MODE == 2) {
                    Cp[(int64_t)r * ld + cc] = -0.5 * x;
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

// ---- patches of dimension <= 16: one wave per patch, every iteration in LDS
//
// The 16 x 16 (identity-padded) patch and its iterates live in LDS with a row
// stride of 17 doubles (a stride of 16 puts the 16 lanes reading one column
// on one bank pair). A 16 x 16 x 16 product is four v_mfma_f64_16x16x4f64.
constexpr int SP = 16;
constexpr int SPL = 17;
constexpr int SPM = SP * SPL;

// acc += A B (16 x 16 operands in LDS); acc[v] = (A B)[4v + g][i]
__device__ __forceinline__ D4 mm16(const double *A, const double *B, int i, int g, D4 acc) {
#pragma unroll
    for (int k4 = 0; k4 < SP; k4 += 4)
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(A[i * SPL + k4 + g], B[(k4 + g) * SPL + i], acc, 0, 0,
                                                   0);
    return acc;
}

__device__ __forceinline__ void st16(double *D, D4 v, int i, int g) {
#pragma unroll
    for (int q = 0; q < 4; ++q) D[(4 * q + g) * SPL + i] = v[q];
}

__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    return x;
}

__device__ __forceinline__ double wave_max(double x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x = fmax(x, __shfl_xor(x, o));
    return x;
}

// The coupled iteration of patch first + blockIdx.x (n = b |C_i| <= 16).
// TANGENT = false: iterate until ||I - Z Y||_F < tol (or maxit updates, or
// divergence), write the factor's block row i (rows of block i of Z / sqrt(c))
// into zout, and the update count and final residual into iters / resid.
// TANGENT = true: iterate the pair and its tangent (Y', Z') seeded with
// Y0' = E, Z0' = 0, where E holds dL/dZ of block row i in the rows of block
// i (dz: column-major blocks on S's pattern); after convergence `extra` more
// updates; write Z' (-> L_f(A/c, E), f(x) = x^{-1/2}) to dP[p][16][16].
// cs[p] = c, the Gershgorin bound.
template <bool TANGENT>
__global__ __launch_bounds__(64) void k_ns_small(int64_t first, int b, const int64_t *__restrict__ row_ptr,
                                                 const int32_t *__restrict__ col_ind,
                                                 const double *__restrict__ svals,
                                                 const double *__restrict__ dz, double tol, int maxit,
                                                 int extra, double *__restrict__ zout,
                                                 double *__restrict__ dP, double *__restrict__ cs,
                                                 int *__restrict__ iters, double *__restrict__ resid) {
    __shared__ double sm[(TANGENT ? 10 : 5) * SPM];
    const int64_t p = blockIdx.x;
    const int64_t row = first + p;
    const int lane = threadIdx.x;
    const int i = lane & 15, g = lane >> 4;
    double *Y = sm, *Yn = sm + SPM, *Zm = sm + 2 * SPM, *Zn = sm + 3 * SPM, *T = sm + 4 * SPM;
    double *Yd = sm + 5 * SPM, *Ydn = sm + 6 * SPM, *Zd = sm + 7 * SPM, *Zdn = sm + 8 * SPM,
           *Td = sm + 9 * SPM;
    const int64_t s0 = row_ptr[row];
    const int nb = (int)(row_ptr[row + 1] - s0);
    const int n = nb * b;
    for (int t = lane; t < SP * SP; t += 64) {
        const int r = t / SP, c = t - (t / SP) * SP;
        const double d = (r == c) ? 1.0 : 0.0;
        Zm[r * SPL + c] = d;
        Y[r * SPL + c] = (r >= n) ? d : 0.0; // padding rows keep the identity
        if (TANGENT) {
            Yd[r * SPL + c] = 0.0;
            Zd[r * SPL + c] = 0.0;
        }
    }
    __syncthreads();
    // lane r < n gathers patch row r and its absolute row sum
    double rsum = 0.0;
    if (lane < n) {
        const int pb = lane / b, rr = lane - pb * b;
        const int32_t ci = col_ind[s0 + pb];
        for (int q = 0; q < nb; ++q) {
            const int64_t sb = find_block(row_ptr, col_ind, ci, col_ind[s0 + q]);
            if (sb < 0) continue;
            const double *blk = svals + sb * (int64_t)b * b; // column-major
            for (int cc = 0; cc < b; ++cc) {
                const double v = blk[cc * b + rr];
                rsum += fabs(v);
                Y[lane * SPL + q * b + cc] = v;
            }
        }
    }
    const double cmax = wave_max(rsum);
    const double c = cmax > 0.0 ? cmax : 1.0;
    if (lane < n)
        for (int cc = 0; cc < n; ++cc) Y[lane * SPL + cc] *= 1.0 / c;
    int me = 0;
    for (int q = 0; q < nb; ++q)
        if (col_ind[s0 + q] == (int32_t)row) me = q;
    if (TANGENT) {
        const int per = b * b;
        for (int t = lane; t < nb * per; t += 64) {
            const int q = t / per, e = t - q * per;
            const int cc = e / b, rr = e - cc * b;
            Yd[(me * b + rr) * SPL + q * b + cc] = dz[(s0 + q) * (int64_t)per + e];
        }
    }
    __syncthreads();
    int it = 0, left = extra;
    bool conv = false;
    double res = 0.0;
    const D4 zero4 = {0, 0, 0, 0};
    for (;;) {
        const D4 zy = mm16(Zm, Y, i, g, zero4);
        double d2 = 0.0;
        D4 t4;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int r = 4 * q + g;
            const double id = (r == i) ? 1.0 : 0.0;
            const double d = id - zy[q];
            d2 = __builtin_fma(d, d, d2);
            t4[q] = 1.5 * id - 0.5 * zy[q];
        }
        res = sqrt(wave_sum(d2));
        if (!conv) {
            if (res < tol) conv = true;
            else if (!(res <= 100.0) || it == maxit) break;
        }
        if (conv) {
            if (left == 0) break;
            --left;
        }
        st16(T, t4, i, g);
        if (TANGENT) {
            const D4 td = mm16(Zm, Yd, i, g, mm16(Zd, Y, i, g, zero4));
            st16(Td, -0.5 * td, i, g);
        }
        __syncthreads();
        st16(Yn, mm16(Y, T, i, g, zero4), i, g);
        st16(Zn, mm16(T, Zm, i, g, zero4), i, g);
        if (TANGENT) {
            st16(Ydn, mm16(Y, Td, i, g, mm16(Yd, T, i, g, zero4)), i, g);
            st16(Zdn, mm16(T, Zd, i, g, mm16(Td, Zm, i, g, zero4)), i, g);
        }
        __syncthreads();
        double *sw = Y;
        Y = Yn;
        Yn = sw;
        sw = Zm;
        Zm = Zn;
        Zn = sw;
        if (TANGENT) {
            sw = Yd;
            Yd = Ydn;
            Ydn = sw;
            sw = Zd;
            Zd = Zdn;
            Zdn = sw;
        }
        ++it;
    }
    if (lane == 0) {
        cs[p] = c;
        if (iters) iters[p] = it;
        if (resid) resid[p] = res;
    }
    if (TANGENT) {
        for (int t = lane; t < SP * SP; t += 64) {
            const int r = t / SP, cc = t - (t / SP) * SP;
            dP[p * (SP * SP) + t] = Zd[r * SPL + cc];
        }
    } else {
        const double sc = 1.0 / sqrt(c);
        const int per = b * b;
        for (int t = lane; t < nb * per; t += 64) {
            const int q = t / per, e = t - q * per;
            const int cc = e / b, rr = e - cc * b;
            zout[(s0 + q) * (int64_t)per + e] = Zm[(me * b + rr) * SPL + q * b + cc] * sc;
        }
    }
}

// Tangent seed of the tiled iteration: Yd = E (dL/dZ of block row i in the
// rows of block i, zero elsewhere), Zd = 0; ld x ld per patch.
__global__ __launch_bounds__(256) void k_patch_seed(int64_t first, int b, const int64_t *__restrict__ row_ptr,
                                                    const int32_t *__restrict__ col_ind,
                                                    const double *__restrict__ dz, int ld,
                                                    double *__restrict__ Yd, double *__restrict__ Zd) {
    const int64_t p = blockIdx.x;
    const int64_t row = first + p;
    const int64_t s0 = row_ptr[row];
    const int nb = (int)(row_ptr[row + 1] - s0);
    double *__restrict__ Yp = Yd + p * (int64_t)ld * ld;
    double *__restrict__ Zp = Zd + p * (int64_t)ld * ld;
    const int64_t ldd = (int64_t)ld * ld;
    for (int64_t t = threadIdx.x; t < ldd; t += blockDim.x) {
        Yp[t] = 0.0;
        Zp[t] = 0.0;
    }
    __syncthreads();
    int me = 0;
    for (int q = 0; q < nb; ++q)
        if (col_ind[s0 + q] == (int32_t)row) me = q;
    const int per = b * b;
    for (int t = threadIdx.x; t < nb * per; t += blockDim.x) {
        const int q = t / per, e = t - q * per;
        const int cc = e / b, rr = e - cc * b;
        Yp[(int64_t)(me * b + rr) * ld + q * b + cc] = dz[(s0 + q) * (int64_t)per + e];
    }
}

// dS block s = (p, q) (column-major element e = cc b + rr) += sum over the
// patches i in [first, first + count) that hold both block p and block q
// (i in C_p and in C_q, the pattern being symmetric) of
// c_i^{-3/2} dP_i[pb b + rr][qb b + cc], pb and qb the positions of p and q
// in C_i; patches in ascending order (a fixed summation order).
__global__ __launch_bounds__(256) void k_patch_grad_gather(int64_t nbr, int b, int64_t nnzb,
                                                           const int64_t *__restrict__ row_ptr,
                                                           const int32_t *__restrict__ col_ind, int64_t first,
                                                           int64_t count, int ld, const double *__restrict__ dP,
                                                           const double *__restrict__ cs,
                                                           double *__restrict__ dS) {
    const int per = b * b;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nnzb * per;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t s = t / per;
        const int e = (int)(t - s * per);
        const int cc = e / b, rr = e - cc * b;
        int64_t lo = 0, hi = nbr - 1;
        while (lo < hi) {
            const int64_t mid = (lo + hi + 1) >> 1;
            if (row_ptr[mid] <= s) lo = mid;
            else hi = mid - 1;
        }
        const int32_t pr = (int32_t)lo, qc = col_ind[s];
        int64_t ia = row_ptr[pr], ea = row_ptr[pr + 1], ib = row_ptr[qc], eb = row_ptr[qc + 1];
        double acc = 0.0;
        while (ia < ea && ib < eb) {
            const int32_t xa = col_ind[ia], xb = col_ind[ib];
            if (xa < xb) {
                ++ia;
            } else if (xb < xa) {
                ++ib;
            } else {
                const int64_t pi = xa;
                if (pi >= first && pi < first + count) {
                    const int64_t l = pi - first;
                    const int pb = (int)(find_block(row_ptr, col_ind, pi, pr) - row_ptr[pi]);
                    const int qb = (int)(find_block(row_ptr, col_ind, pi, qc) - row_ptr[pi]);
                    const double w = 1.0 / (cs[l] * sqrt(cs[l]));
                    acc += w * dP[l * (int64_t)ld * ld + (int64_t)(pb * b + rr) * ld + qb * b + cc];
                }
                ++ia;
                ++ib;
            }
        }
        dS[t] += acc;
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
    return launch_ns_gemm2(mode, count, ld, A, B, nullptr, nullptr, C, resid, st);
}

hipError_t launch_ns_gemm2(int mode, int64_t count, int ld, const double *A, const double *B,
                           const double *A2, const double *B2, double *C, double *resid,
                           hipStream_t st) {
    const int tiles = ld / 64;
    const dim3 grid((unsigned)(count * tiles * tiles));
    switch (mode) {
    case 0: hipLaunchKernelGGL(k_ns_gemm<0>, grid, dim3(256), 0, st, ld, A, B, A2, B2, C, resid); break;
    case 1: hipLaunchKernelGGL(k_ns_gemm<1>, grid, dim3(256), 0, st, ld, A, B, A2, B2, C, resid); break;
    case 2: hipLaunchKernelGGL(k_ns_gemm<2>, grid, dim3(256), 0, st, ld, A, B, A2, B2, C, resid); break;
    case 3: hipLaunchKernelGGL(k_ns_gemm<3>, grid, dim3(256), 0, st, ld, A, B, A2, B2, C, resid); break;
    default: return hipErrorInvalidValue;
    }
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


int ns_small_max() { return SP; }

hipError_t launch_ns_small(int64_t first, int64_t count, int b, const int64_t *row_ptr, const int32_t *col_ind,
                           const double *svals, double tol, int maxit, double *zout, double *cs, int *iters,
                           double *resid, hipStream_t st) {
    hipLaunchKernelGGL(k_ns_small<false>, dim3((unsigned)count), dim3(64), 0, st, first, b, row_ptr, col_ind,
                       svals, (const double *)nullptr, tol, maxit, 0, zout, (double *)nullptr, cs, iters, resid);
    return hipGetLastError();
}

hipError_t launch_ns_small_tangent(int64_t first, int64_t count, int b, const int64_t *row_ptr,
                                   const int32_t *col_ind, const double *svals, const double *dz, double tol,
                                   int maxit, int extra, double *dP, double *cs, int *iters, double *resid,
                                   hipStream_t st) {
    hipLaunchKernelGGL(k_ns_small<true>, dim3((unsigned)count), dim3(64), 0, st, first, b, row_ptr, col_ind,
                       svals, dz, tol, maxit, extra, (double *)nullptr, dP, cs, iters, resid);
    return hipGetLastError();
}

hipError_t launch_patch_seed(int64_t first, int64_t count, int b, const int64_t *row_ptr, const int32_t *col_ind,
                             const double *dz, int ld, double *Yd, double *Zd, hipStream_t st) {
    hipLaunchKernelGGL(k_patch_seed, dim3((unsigned)count), dim3(256), 0, st, first, b, row_ptr, col_ind, dz, ld,
                       Yd, Zd);
    return hipGetLastError();
}

hipError_t launch_patch_grad_gather(int64_t nbr, int b, int64_t nnzb, const int64_t *row_ptr,
                                    const int32_t *col_ind, int64_t first, int64_t count, int ld,
                                    const double *dP, const double *cs, double *dS, hipStream_t st) {
    int64_t g = (nnzb * b * b + 255) / 256;
    if (g > 16384) g = 16384;
    if (g < 1) g = 1;
    hipLaunchKernelGGL(k_patch_grad_gather, dim3((unsigned)g), dim3(256), 0, st, nbr, b, nnzb, row_ptr, col_ind,
                       first, count, ld, dP, cs, dS);
    return hipGetLastError();
}

} // namespace rg
