// Group-resolved dots of the projected-DOS moments.
//
// P_g is diagonal, so <a|P_g|b> is conj(a) . b summed over the rows of group
// g and all sample columns. Rows arrive sorted by group (a CSR of groups over
// rows); each group's row list is cut into chunks of about kChunkElems
// (row, column) elements, one workgroup per chunk writes its partial dots,
// and one wave per group sums its chunks in a fixed order, so the result is
// the same from run to run.
#include "rg_internal.h"

namespace rg {
namespace {

constexpr int kThreads = 256;

template <typename S> struct Cplx;
template <> struct Cplx<double> { using T = double2; };
template <> struct Cplx<float> { using T = float2; };

// conj(a) b accumulated in FP64 (FP32 products are exact in FP64)
__device__ __forceinline__ void cdot(double2 &acc, double ax, double ay, double bx, double by) {
    acc.x = __builtin_fma(ax, bx, __builtin_fma(ay, by, acc.x));
    acc.y = __builtin_fma(ax, by, __builtin_fma(-ay, bx, acc.y));
}

template <typename S, bool CV>
__device__ __forceinline__ void load_elem(const void *v, int64_t idx, double &x, double &y) {
    if constexpr (CV) {
        const auto t = static_cast<const typename Cplx<S>::T *>(v)[idx];
        x = (double)t.x;
        y = (double)t.y;
    } else {
        x = (double)static_cast<const S *>(v)[idx];
        y = 0.0;
    }
}

__device__ __forceinline__ double2 wave_sum(double2 v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        v.x += __shfl_xor(v.x, off);
        v.y += __shfl_xor(v.y, off);
    }
    return v;
}

// ndot 1: <v0|v1>; 2: <v0|v0>, <v0|v1>; 3: <v0|v0>, <v0|v1>, <v1|v1>.
// partial[c * 3 + d] for chunk c; chunk c covers rows[chunk_ptr[c] .. chunk_ptr[c+1]).
template <typename S, bool CV>
__global__ __launch_bounds__(kThreads) void k_group_dots(
    const int64_t *__restrict__ chunk_ptr, const int32_t *__restrict__ rows, int R, int ndot,
    const void *__restrict__ v0, const void *__restrict__ v1, double2 *__restrict__ partial) {
    const int64_t c = blockIdx.x;
    const int64_t r0 = chunk_ptr[c];
    const int64_t nel = (chunk_ptr[c + 1] - r0) * R;
    double2 d0{0.0, 0.0}, d1{0.0, 0.0}, d2{0.0, 0.0};
    for (int64_t t = threadIdx.x; t < nel; t += kThreads) {
        const int64_t rl = t / R;
        const int col = (int)(t - rl * R);
        const int64_t idx = (int64_t)rows[r0 + rl] * R + col;
        double ax, ay, bx, by;
        load_elem<S, CV>(v0, idx, ax, ay);
        load_elem<S, CV>(v1, idx, bx, by);
        if (ndot == 1) {
            cdot(d0, ax, ay, bx, by);
        } else {
            cdot(d0, ax, ay, ax, ay);
            cdot(d1, ax, ay, bx, by);
            if (ndot == 3) cdot(d2, bx, by, bx, by);
        }
    }
    __shared__ double2 sm[kThreads / 64][3];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    d0 = wave_sum(d0);
    d1 = wave_sum(d1);
    d2 = wave_sum(d2);
    if (lane == 0) {
        sm[w][0] = d0;
        sm[w][1] = d1;
        sm[w][2] = d2;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        double2 s{0.0, 0.0};
#pragma unroll
        for (int q = 0; q < kThreads / 64; ++q) {
            s.x += sm[q][threadIdx.x].x;
            s.y += sm[q][threadIdx.x].y;
        }
        partial[c * 3 + threadIdx.x] = s;
    }
}

// out[d * G + g] = sum of partial[c * 3 + d] over group g's chunks, d < ndot.
__global__ __launch_bounds__(kThreads) void k_group_finish(const int64_t *__restrict__ grp_chunk,
                                                           int64_t G, int ndot,
                                                           const double2 *__restrict__ partial,
                                                           double2 *__restrict__ out) {
    const int64_t g = (int64_t)blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (g >= G) return;
    const int64_t c0 = grp_chunk[g], c1 = grp_chunk[g + 1];
    for (int d = 0; d < ndot; ++d) {
        double2 s{0.0, 0.0};
        for (int64_t c = c0 + lane; c < c1; c += 64) {
            const double2 p = partial[c * 3 + d];
            s.x += p.x;
            s.y += p.y;
        }
        s = wave_sum(s);
        if (lane == 0) out[d * G + g] = s;
    }
}

} // namespace

hipError_t launch_group_dots(int64_t nchunks, const int64_t *chunk_ptr, const int32_t *rows,
                             int R, int ndot, VecMode vm, const void *v0, const void *v1,
                             double2 *partial, hipStream_t st) {
    if (nchunks <= 0) return hipSuccess;
    const dim3 grid((unsigned)nchunks), block(kThreads);
    if (vm.fp32) {
        if (vm.complex_vec)
            hipLaunchKernelGGL((k_group_dots<float, true>), grid, block, 0, st, chunk_ptr, rows, R,
                               ndot, v0, v1, partial);
        else
            hipLaunchKernelGGL((k_group_dots<float, false>), grid, block, 0, st, chunk_ptr, rows,
                               R, ndot, v0, v1, partial);
    } else {
        if (vm.complex_vec)
            hipLaunchKernelGGL((k_group_dots<double, true>), grid, block, 0, st, chunk_ptr, rows,
                               R, ndot, v0, v1, partial);
        else
            hipLaunchKernelGGL((k_group_dots<double, false>), grid, block, 0, st, chunk_ptr, rows,
                               R, ndot, v0, v1, partial);
    }
    return hipGetLastError();
}

hipError_t launch_group_finish(int64_t G, const int64_t *grp_chunk, int ndot,
                               const double2 *partial, double2 *out, hipStream_t st) {
    if (G <= 0) return hipSuccess;
    const int64_t wpb = kThreads / 64;
    hipLaunchKernelGGL(k_group_finish, dim3((unsigned)((G + wpb - 1) / wpb)), dim3(kThreads), 0,
                       st, grp_chunk, G, ndot, partial, out);
    return hipGetLastError();
}

} // namespace rg
