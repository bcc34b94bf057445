// Slater-Koster assembly of H or S on the device (rsatb/model/sktb/
// sktb_construction.hpp:327-500, the reference's per-edge block evaluation).
//
// Two kernels:
//   k_sk_edges   one wave per edge: evaluates the edge's b x b block
//                (row-major) from the bond vector (the table pair of the
//                edge's atom types if its cutoff covers the bond, else the
//                transposed block of the reversed pair at -vector,
//                :476-500; per channel the radial value times the two shell
//                rotations of the aligned sigma/pi/delta entries) and marks
//                edges with an entry above 1e-14 (the reference stores only
//                those, :385);
// then, once the host has laid out the pattern of the kept blocks,
//   k_sk_gather       sums each stored block's sources in a fixed order (the
//                     onsite diagonal, then its edges in ascending order,
//                     times their k phases, or times a bond component for the
//                     velocity factors) into the column-major device layout.
// All arithmetic is FP64; this is setup work, bound by the edge count.
#include "rg_internal.h"

#include <algorithm>

namespace rg {
namespace {

// canonical real d quadrupoles in magnetic order dxy, dyz, dz2, dzx, dx2-y2
// (sktb_construction.hpp:93-115), row-major 3 x 3 each
__constant__ double kCanonD[5][9];

const double kS2 = 0.70710678118654752440; // 1 / sqrt(2)
const double kS6 = 0.40824829046386301637; // 1 / sqrt(6)
const double kCanonDHost[5][9] = {
    {0, kS2, 0, kS2, 0, 0, 0, 0, 0},
    {0, 0, 0, 0, 0, kS2, 0, kS2, 0},
    {-kS6, 0, 0, 0, -kS6, 0, 0, 0, 2 * kS6},
    {0, 0, kS2, 0, 0, 0, kS2, 0, 0},
    {kS2, 0, 0, 0, -kS2, 0, 0, 0, 0},
};

// F[i][j]: column j is ex, ey, ez of the bond frame (:117-131)
struct Frame {
    double f[3][3];
};

__device__ Frame bond_frame(double x, double y, double z) {
    Frame F;
    const double n = sqrt(x * x + y * y + z * z);
    double ez[3] = {0, 0, 0};
    if (n > 1e-12) {
        ez[0] = x / n;
        ez[1] = y / n;
        ez[2] = z / n;
    }
    double ex[3] = {1, 0, 0};
    if (fabs(ez[2]) < 1.0 - 1e-12) {
        // (0, 0, 1) x ez, normalised
        const double cx = -ez[1], cy = ez[0];
        const double cn = sqrt(cx * cx + cy * cy);
        if (cn > 1e-12) {
            ex[0] = cx / cn;
            ex[1] = cy / cn;
        } else {
            ex[0] = 0.0;
            ex[1] = 0.0;
        }
        ex[2] = 0.0;
    }
    const double ey[3] = {ez[1] * ex[2] - ez[2] * ex[1], ez[2] * ex[0] - ez[0] * ex[2],
                          ez[0] * ex[1] - ez[1] * ex[0]};
    for (int i = 0; i < 3; ++i) {
        F.f[i][0] = ex[i];
        F.f[i][1] = ey[i];
        F.f[i][2] = ez[i];
    }
    return F;
}

// Shell rotation entry R_l[loc][glo] (:140-177): p in magnetic order (py, pz,
// px) maps to the axes (y, z, x); d is <C_glo, F C_loc F^T>_Frobenius.
__device__ double shell_rot(int l, int loc, int glo, const Frame &F) {
    if (l == 0) return 1.0;
    if (l == 1) {
        const int ax[3] = {1, 2, 0};
        return F.f[ax[glo]][ax[loc]];
    }
    const double *cl = kCanonD[loc];
    const double *cg = kCanonD[glo];
    double acc = 0.0;
    for (int p = 0; p < 3; ++p)
        for (int q = 0; q < 3; ++q) {
            // (F C_loc F^T)[p][q]
            double m = 0.0;
            for (int r = 0; r < 3; ++r)
                for (int s = 0; s < 3; ++s) m += F.f[p][r] * cl[r * 3 + s] * F.f[q][s];
            acc += cg[p * 3 + q] * m;
        }
    return acc;
}

// Radial value of channel c at distance d (radial.hpp:139-195)
__device__ double radial(const SkDev &t, int c, double d) {
    if (d < 0.0) return 0.0;
    const double rcut = t.ch_r_cut[c];
    if (d > rcut) return 0.0;
    const int off = t.ch_off[c], n = t.ch_n[c];
    const double *y = t.rad_y + off;
    if (t.ch_kind[c] == 0) { // natural cubic spline
        const double *x = t.rad_x + off;
        const double *y2 = t.rad_y2 + off;
        if (d <= x[0]) return y[0];
        // first knot >= d
        int lo = 0, hi = n;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (x[mid] < d) lo = mid + 1;
            else hi = mid;
        }
        if (lo == n) return y[n - 1];
        const int a0 = lo - 1, a1 = lo;
        const double h = x[a1] - x[a0];
        if (h <= 1e-12) return y[a1];
        const double a = (x[a1] - d) / h, b = (d - x[a0]) / h;
        return a * y[a0] + b * y[a1] + ((a * a * a - a) * y2[a0] + (b * b * b - b) * y2[a1]) * (h * h) / 6.0;
    }
    // exp_poly with a cosine taper past r_on
    double w = 1.0;
    const double ron = t.ch_r_on[c];
    if (d > ron) {
        const double span = rcut - ron;
        w = span <= 1e-12 ? 0.0 : 0.5 * (1.0 + cos(M_PI * (d - ron) / span));
    }
    double poly = 0.0, pw = 1.0;
    for (int k = 0; k < n; ++k) {
        poly += y[k] * pw;
        pw *= d;
    }
    return w * exp(-t.ch_lambda[c] * d) * poly;
}

// One wave per edge (grid-stride). The wave first stages what every element
// of the block shares: the 9 p and 25 d rotation entries of the bond frame
// (lanes 0-33) and the radial value of each channel of the table pair (lanes
// 34 on); then each lane sums the channel terms of its elements from LDS.
constexpr int kSkWaves = 4;
constexpr int kSkMaxPairCh = 30; // channels of one ordered type pair (s, p, d: at most 14)

__global__ __launch_bounds__(64 * kSkWaves) void k_sk_edges(SkDev t, int64_t n_edges,
                                                            const int32_t *__restrict__ type,
                                                            const double *__restrict__ pos, SkCell cell,
                                                            const int32_t *__restrict__ edge_index,
                                                            const int32_t *__restrict__ edge_shift,
                                                            double *__restrict__ eblk,
                                                            uint8_t *__restrict__ flags) {
    __shared__ double s_r1[kSkWaves][9], s_r2[kSkWaves][25], s_rad[kSkWaves][kSkMaxPairCh];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int b = t.norb, bb = b * b, T = t.n_types;
    for (int64_t e = (int64_t)blockIdx.x * kSkWaves + w; e < n_edges; e += (int64_t)gridDim.x * kSkWaves) {
        const int i = edge_index[2 * e], j = edge_index[2 * e + 1];
        const double s0 = edge_shift[3 * e], s1 = edge_shift[3 * e + 1], s2 = edge_shift[3 * e + 2];
        double v[3];
        for (int q = 0; q < 3; ++q)
            v[q] = pos[3 * j + q] + s0 * cell.c[0][q] + s1 * cell.c[1][q] + s2 * cell.c[2][q] - pos[3 * i + q];
        const double d = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        int rt = type[i], ct = type[j];
        bool swap = false;
        bool ok = d > 1e-12 && d <= t.pair_cutoff[rt * T + ct] + 1e-12;
        if (!ok && d > 1e-12 && d <= t.pair_cutoff[ct * T + rt] + 1e-12) {
            // reversed pair at -v, transposed back (sktb_construction.hpp:476-500)
            swap = true;
            const int tt = rt;
            rt = ct;
            ct = tt;
            v[0] = -v[0];
            v[1] = -v[1];
            v[2] = -v[2];
            ok = true;
        }
        double *out = eblk + e * bb;
        if (!ok) {
            for (int el = lane; el < bb; el += 64) out[el] = 0.0;
            if (lane == 0) flags[e] = 0;
            continue;
        }
        const int c0 = t.pair_ch[2 * (rt * T + ct)], nch = t.pair_ch[2 * (rt * T + ct) + 1] - c0;
        {
            const Frame F = bond_frame(v[0], v[1], v[2]);
            if (lane < 9) {
                s_r1[w][lane] = shell_rot(1, lane / 3, lane % 3, F);
            } else if (lane < 34) {
                const int q = lane - 9;
                s_r2[w][q] = shell_rot(2, q / 5, q % 5, F);
            } else if (lane - 34 < nch) {
                const int ch = c0 + lane - 34;
                s_rad[w][lane - 34] = d > t.ch_r_cut[ch] + 1e-12 ? 0.0 : radial(t, ch, d);
            }
        }
        __builtin_amdgcn_wave_barrier(); // LDS operations of one wave complete in order
        bool nz = false;
        for (int el = lane; el < bb; el += 64) {
            const int a = el / b, c = el - a * b;
            const int ra = swap ? c : a, ca = swap ? a : c;
            const int la = t.orb_l[rt * b + ra], ma = t.orb_m[rt * b + ra];
            const int lc = t.orb_l[ct * b + ca], mc = t.orb_m[ct * b + ca];
            const double sign = (la > lc && ((la + lc) & 1)) ? -1.0 : 1.0;
            auto rot = [&](int l, int loc, int glo) -> double {
                return l == 0 ? 1.0 : (l == 1 ? s_r1[w][loc * 3 + glo] : s_r2[w][loc * 5 + glo]);
            };
            double value = 0.0;
            for (int k = 0; k < nch; ++k) {
                const int ch = c0 + k;
                if (t.ch_row_l[ch] != la || t.ch_col_l[ch] != lc) continue;
                const int m = t.ch_bond[ch];
                const double r = m == 0 ? rot(la, la, ma) * rot(lc, lc, mc)
                                        : rot(la, la - m, ma) * rot(lc, lc - m, mc) +
                                              rot(la, la + m, ma) * rot(lc, lc + m, mc);
                value += sign * s_rad[w][k] * r;
            }
            out[el] = value;
            nz = nz || fabs(value) > 1e-14;
        }
        const uint64_t any = __ballot(nz);
        if (lane == 0) flags[e] = any != 0;
        __builtin_amdgcn_wave_barrier(); // staged values are read before the next edge's writes
    }
}

__global__ __launch_bounds__(256) void k_sk_gather(int64_t nnzb, int b, const int64_t *__restrict__ src_ptr,
                                                   const int32_t *__restrict__ src, const double2 *__restrict__ phase,
                                                   const int32_t *__restrict__ row_of, const int32_t *__restrict__ type,
                                                   const double *__restrict__ onsite, const double *__restrict__ eblk,
                                                   double *__restrict__ vals, int mode) {
    const int bb = b * b;
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= nnzb * bb) return;
    const int64_t s = tid / bb;
    const int el = (int)(tid - s * bb), i = el / b, j = el - i * b;
    double re = 0.0, im = 0.0;
    for (int64_t q = src_ptr[s]; q < src_ptr[s + 1]; ++q) {
        const int e = src[q];
        double x;
        if (e < 0) // onsite diagonal of the row's atom
            x = i == j ? onsite[type[row_of[s]] * b + i] : 0.0;
        else
            x = eblk[(int64_t)e * bb + el];
        if (mode == 1) { // k phase
            const double2 ph = phase[q];
            re += x * ph.x;
            im += x * ph.y;
        } else if (mode == 2) { // real weight (a bond component)
            re += x * phase[q].x;
        } else {
            re += x;
        }
    }
    // column-major block: element (i, j) at j b + i
    const int64_t o = s * bb + (int64_t)j * b + i;
    if (mode == 1) {
        vals[2 * o] = re;
        vals[2 * o + 1] = im;
    } else {
        vals[o] = re;
    }
}

} // namespace

hipError_t sk_upload_constants() {
    static bool done = false;
    if (done) return hipSuccess;
    hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(kCanonD), kCanonDHost, sizeof(kCanonDHost));
    if (e == hipSuccess) done = true;
    return e;
}

hipError_t launch_sk_edge_blocks(const SkDev &t, int64_t n_edges, const int32_t *type, const double *pos,
                                 const SkCell &cell, const int32_t *edge_index, const int32_t *edge_shift,
                                 double *eblk, uint8_t *flags, hipStream_t st) {
    if (n_edges == 0) return hipSuccess;
    const int64_t grid = std::min<int64_t>((n_edges + kSkWaves - 1) / kSkWaves, 16384);
    hipLaunchKernelGGL(k_sk_edges, dim3((unsigned)grid), dim3(64 * kSkWaves), 0, st, t, n_edges, type, pos, cell,
                       edge_index, edge_shift, eblk, flags);
    return hipGetLastError();
}

int sk_max_pair_channels() { return kSkMaxPairCh; }

hipError_t launch_sk_gather(int64_t nnzb, int b, const int64_t *src_ptr, const int32_t *src, const double2 *phase,
                            const int32_t *row_of, const int32_t *type, const double *onsite, const double *eblk,
                            double *vals, int mode, hipStream_t st) {
    const int64_t n = nnzb * (int64_t)b * b;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_sk_gather, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, nnzb, b, src_ptr, src,
                       phase, row_of, type, onsite, eblk, vals, mode);
    return hipGetLastError();
}

} // namespace rg
