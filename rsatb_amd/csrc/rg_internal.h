// Internal declarations shared by the kernel and engine translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rg {

// Spare (zero) entries after the MFMA SpMM's gather table: a wave fetches up
// to 2 PD - 2 steps of 4 entries past its row, and reads the entries of the
// step after its last fetch (PD <= 6: at most 8 PD - 2 = 46 entries).
constexpr int kXrowSpare = 64;

// Device view of one BSR operand. Values are real (vals) or complex
// interleaved (vals with is_complex), in double or float storage.
struct BsrDev {
    int64_t nbr = 0;
    int32_t b = 0;
    int64_t nnzb = 0;
    const int64_t *row_ptr = nullptr;
    const int32_t *col_ind = nullptr;
    const void *vals = nullptr; // double / double2 / float / float2
    bool is_complex = false;
    // gather table of real b >= 8 matrices, or null: xrow[s * b + kk] =
    // col_ind[s] * b + kk, the X row of column kk of block s (kXrowSpare spare entries)
    const int32_t *xrow = nullptr;
    // processing order of the block rows (MFMA SpMM), or null for stored
    // order: position p of [0, nbr) computes block row sched[p]. The XCD
    // ranges of a launch are whole regions of sreg consecutive positions.
    const int32_t *sched = nullptr;
    int32_t sreg = 0;
    // > 0 when every block row holds exactly ulen blocks and row_ptr[0] = 0:
    // row br's blocks are then br * ulen .. br * ulen + ulen - 1, and the VALU
    // SpMM computes them instead of loading row_ptr
    int32_t ulen = 0;
};

// Epilogue of a SpMM launch. y = sum_blocks A x is finished as
//   EPI_PLAIN      Y = y
//   EPI_FIRST      Y = alpha y + beta Xs;           dots <Xs|Xs>, <Xs|Y>, <Y|Y>
//   EPI_STEP       Y = alpha y + beta Xs - Xp;      dots <Y|Y>, <Y|Xs>
//   EPI_PLAINSTEP  Y = alpha y + beta Xs - Xp;      dot  <Psi|Y>
//   EPI_ADJ        Y = alpha y + beta Xs - Xp + cm Qm + cz Qz + cp Qp (no dots):
//                  one step of the adjoint recurrence, the Q vectors being the
//                  forward states psi_{n-1}, psi_n, psi_{n+1}
// Dots are per sample column, conjugate-linear in the left argument, and
// leave one (re, im) FP64 partial per workgroup and column in `partials`
// laid out [gridDim.x][ndot][R].
enum Epi { EPI_PLAIN = 0, EPI_FIRST = 1, EPI_STEP = 2, EPI_PLAINSTEP = 3, EPI_ADJ = 4 };

__host__ __device__ constexpr inline int epi_ndot(int mode) {
    return mode == EPI_FIRST ? 3 : mode == EPI_STEP ? 2 : mode == EPI_PLAINSTEP ? 1 : 0;
}

struct EpiArgs {
    int mode = EPI_PLAIN;
    double alpha = 1.0, beta = 0.0;
    const void *xs = nullptr;
    const void *xp = nullptr;
    const void *psi = nullptr;
    double2 *partials = nullptr;
    const void *qm = nullptr, *qz = nullptr, *qp = nullptr; // EPI_ADJ
    double cm = 0.0, cz = 0.0, cp = 0.0;
};

// Vector storage of the moment loop.
struct VecMode {
    bool complex_vec = false; // [rows][R] double2 instead of double
    bool fp32 = false;        // float storage (precision form)
};

// Number of workgroups the block-row-gather (VALU) SpMM uses for nbr block
// rows and R columns; spmm_max_grid() bounds every SpMM launch's grid (the
// dot-partial buffer is sized by it).
int spmm_grid(int64_t nbr, int R);
int spmm_max_grid();
// XCD-contiguous row ranges in the MFMA SpMM (default on).
void set_xcd_map(int on);
void set_mfma_ntcap(int nt);
// cap on the MFMA SpMM's workgroups (a multiple of 8, <= spmm_max_grid());
// anything else: two workgroups per compute unit (the default)
void set_mfma_grid_cap(int g);
void set_b1_wide(int mode);
void set_uniform_rows(int on);
// bytes of a row one lane owns in the b = 1 slab SpMM: 16, 32, 64, 0 (off);
// any other value restores the default
void set_b1_slab(int bytes);
// fewest workgroups of a slab launch (smaller problems take the other
// layouts); negative: the default
void set_b1_slab_min_grid(int g);
// sample columns a lane (1, 2 or 4; anything else: 2) in the gradient
// accumulation for 8 < R <= 32
void set_grad_wide(int cols);
void set_mfma_pd32(int p);
void set_mfma_pd64(int p);
// prefetch depth (2 .. 6) of the sixteen-tile complex MFMA SpMM for one
// precision and kind (plain product or step)
void set_mfma_pd16(int fp32, int step, int p);

// Device matrices hold every b x b block column-major; uploads arrive
// row-major (rsatb's layout) and are transposed on the device.
hipError_t launch_block_transpose(const double *in, double *out, int64_t nnzb, int b, bool cplx,
                                  hipStream_t stream);

// The grid (workgroups, so the number of dot partials) launch_spmm uses
// for this operand, vector storage and R.
int spmm_launch_grid(const BsrDev &A, VecMode vm, int R);

// Y = epilogue(A X). X, Y: [nbr*b][R] in the storage `vm` describes.
hipError_t launch_spmm(const BsrDev &A, VecMode vm, const void *X, void *Y, int R,
                       const EpiArgs &e, hipStream_t stream, int *grid_out);

// out[d*R + u] = sum_g partials[(g*ndot + d)*R + u], fixed order.
hipError_t launch_reduce(const double2 *partials, int ngroups, int ndot, int R, double2 *out,
                         hipStream_t stream);

// The same for nsteps steps in one launch: step s reduces partials +
// s pstride into out + s ostride (double2 strides).
hipError_t launch_reduce_steps(const double2 *partials, int64_t pstride, int ngroups, int ndot, int R,
                               double2 *out, int64_t ostride, int nsteps, hipStream_t stream);

// Sample block: Rademacher (real or complex with zero imaginary part) or
// random phase (complex). scale = 1/sqrt(rows) computed on the host.
hipError_t launch_samples(int kind, uint64_t seed, int64_t rows, int R, double scale,
                          bool complex_vec, void *out, hipStream_t stream);

// out[p * m + e] = in[perm[p] * m + e] for blocks p < nblk of m scalars.
hipError_t launch_block_perm(const double *in, const int32_t *perm, double *out, int64_t nblk, int64_t m,
                             hipStream_t stream);

// Elementwise conversion double -> float over n scalars.
hipError_t launch_f64_to_f32(const double *in, float *out, int64_t n, hipStream_t stream);

// Copies n FP64 samples (real, or complex interleaved) into loop storage:
// widening real to complex (zero imaginary part) and/or narrowing to FP32.
hipError_t launch_stage(const double *in, bool in_complex, void *out, bool out_complex,
                        bool out_fp32, int64_t n, hipStream_t stream);

// ---- projected moments (rg_pdos.hip) ----
// partial[c][3] for chunk c = rows[chunk_ptr[c] .. chunk_ptr[c+1]): ndot 1:
// <v0|v1>; 2: <v0|v0>, <v0|v1>; 3: <v0|v0>, <v0|v1>, <v1|v1>, summed over the
// chunk's rows and all R columns.
hipError_t launch_group_dots(int64_t nchunks, const int64_t *chunk_ptr, const int32_t *rows,
                             int R, int ndot, VecMode vm, const void *v0, const void *v1,
                             double2 *partial, hipStream_t st);
// out[d * G + g] = sum of group g's chunk partials (chunks grp_chunk[g] ..
// grp_chunk[g+1]), fixed order.
hipError_t launch_group_finish(int64_t G, const int64_t *grp_chunk, int ndot,
                               const double2 *partial, double2 *out, hipStream_t st);

// ---- graph-local S^{-1/2} patches (rg_isqrt.hip) ----
hipError_t launch_patch_gather(int64_t first, int64_t count, int b, const int64_t *row_ptr,
                               const int32_t *col_ind, const double *vals, int ld, double *Y,
                               double *Z, double *cs, hipStream_t st);
// mode 0: C = 1.5 I - 0.5 A B and resid[p] += ||I - A B||_F^2; mode 1: C = A B
hipError_t launch_ns_gemm(int mode, int64_t count, int ld, const double *A, const double *B,
                          double *C, double *resid, hipStream_t st);
// the tiled products plus the tangent modes: 2: C = -0.5 (A B + A2 B2),
// 3: C = A B + A2 B2
hipError_t launch_ns_gemm2(int mode, int64_t count, int ld, const double *A, const double *B,
                           const double *A2, const double *B2, double *C, double *resid,
                           hipStream_t st);
// largest patch dimension of the one-wave kernels (16)
int ns_small_max();
// patches of dimension <= 16: the whole iteration per patch in one launch;
// the factor's block rows into zout (column-major blocks on S's pattern),
// cs / iters / resid per patch
hipError_t launch_ns_small(int64_t first, int64_t count, int b, const int64_t *row_ptr, const int32_t *col_ind,
                           const double *svals, double tol, int maxit, double *zout, double *cs, int *iters,
                           double *resid, hipStream_t st);
// the tangent of the same iteration seeded with dL/dZ (dz, on S's pattern):
// dP[p][16][16] -> L_f(A/c, E), f(x) = x^{-1/2}, after `extra` updates past
// convergence
hipError_t launch_ns_small_tangent(int64_t first, int64_t count, int b, const int64_t *row_ptr,
                                   const int32_t *col_ind, const double *svals, const double *dz, double tol,
                                   int maxit, int extra, double *dP, double *cs, int *iters, double *resid,
                                   hipStream_t st);
// tiled tangent seed: Yd = dL/dZ rows of block i, Zd = 0
hipError_t launch_patch_seed(int64_t first, int64_t count, int b, const int64_t *row_ptr, const int32_t *col_ind,
                             const double *dz, int ld, double *Yd, double *Zd, hipStream_t st);
// dS (column-major blocks on S's pattern) += the patch derivatives dP of
// patches [first, first + count), each scaled by c^{-3/2}
hipError_t launch_patch_grad_gather(int64_t nbr, int b, int64_t nnzb, const int64_t *row_ptr,
                                    const int32_t *col_ind, int64_t first, int64_t count, int ld,
                                    const double *dP, const double *cs, double *dS, hipStream_t st);
hipError_t launch_patch_extract(int64_t first, int64_t count, int b, const int64_t *row_ptr,
                                const int32_t *col_ind, int ld, const double *Z, const double *cs,
                                double *out, hipStream_t st);
hipError_t launch_bsr_ct_vals(int64_t nbr, int b, bool cplx, int64_t nnzb, const int64_t *row_ptr,
                              const int32_t *col_ind, const double *in, double *out,
                              hipStream_t st);

// ---- Lanczos vector kernels (rg_lanczos.hip) ----
int lanczos_grid(int64_t rows, int R);
hipError_t launch_lanczos_axpy(int64_t rows, int R, bool cplx, double *w, const double *v,
                               const double *vprev, const double2 *alpha, const double2 *bsq,
                               double2 *partials, int grid, hipStream_t st);
hipError_t launch_lanczos_scale(int64_t rows, int R, bool cplx, const double *w,
                                const double2 *bsq, double tiny, double *v, hipStream_t st);

// ---- adjoint (rg_adjoint.hip) ----
// G[e] += coef * sum_c Re(U[i, c] conj(V[j, c])) for every stored entry e =
// (i, j) of the pattern (row_ptr, col_ind, b); G in column-major block layout.
hipError_t launch_grad_accum(int64_t nbr, int b, int64_t nnzb, const int64_t *row_ptr,
                             const int32_t *col_ind, const double *U, const double *V, int R,
                             bool cplx, double coef, double *G, hipStream_t st);
// the same with a second pair in one pass: G[e] += coef * sum_c (Re(U[i, c]
// conj(V[j, c])) + Re(U2[i, c] conj(V2[j, c])))
hipError_t launch_grad_accum2(int64_t nbr, int b, int64_t nnzb, const int64_t *row_ptr,
                              const int32_t *col_ind, const double *U, const double *V, const double *U2,
                              const double *V2, int R, bool cplx, double coef, double *G, hipStream_t st);
// y = 0 over n doubles
hipError_t launch_zero(double *y, int64_t n, hipStream_t st);
// out[c * rows + r] = (double) in[c * ld_in + r] for r < rows, c < cols
hipError_t launch_widen_cols(const float *in, int64_t ld_in, double *out, int64_t rows, int cols, hipStream_t st);
// y = alpha x + beta y over n doubles
hipError_t launch_axpby(double *y, const double *x, double alpha, double beta, int64_t n, hipStream_t st);

// ---- Slater-Koster assembly (rg_sk.hip) ----
// Device copy of one matrix's SK table; channels grouped by (row type, col
// type), pair_ch[2 (rt T + ct)] .. [+1] the range of a pair's channels.
struct SkDev {
    int n_types = 0, norb = 0, n_channels = 0;
    const int32_t *orb_l = nullptr, *orb_m = nullptr; // [T * norb]
    const double *onsite = nullptr;                   // [T * norb]
    const double *pair_cutoff = nullptr;              // [T * T]
    const int32_t *pair_ch = nullptr;                 // [T * T * 2]
    const int32_t *ch_row_l = nullptr, *ch_col_l = nullptr, *ch_bond = nullptr, *ch_kind = nullptr;
    const int32_t *ch_off = nullptr, *ch_n = nullptr;
    const double *ch_r_cut = nullptr, *ch_lambda = nullptr, *ch_r_on = nullptr;
    const double *rad_x = nullptr, *rad_y = nullptr, *rad_y2 = nullptr;
};
// lattice vectors as rows
struct SkCell {
    double c[3][3];
};
hipError_t sk_upload_constants();
// most channels one ordered type pair may carry on the device
int sk_max_pair_channels();
// every edge's row-major block into eblk[e b^2 ..], and flags[e] = any
// |entry| > 1e-14
hipError_t launch_sk_edge_blocks(const SkDev &t, int64_t n_edges, const int32_t *type, const double *pos,
                                 const SkCell &cell, const int32_t *edge_index, const int32_t *edge_shift,
                                 double *eblk, uint8_t *flags, hipStream_t st);
// stored block s = sum over q in [src_ptr[s], src_ptr[s+1]) of source src[q]
// (-1: the onsite diagonal of row_of[s]'s atom, else edge src[q]); mode 0
// plain sum, 1 times the complex phase[q] (output interleaved re, im), 2
// times the real weight phase[q].x; written column-major
hipError_t launch_sk_gather(int64_t nnzb, int b, const int64_t *src_ptr, const int32_t *src,
                            const double2 *phase, const int32_t *row_of, const int32_t *type,
                            const double *onsite, const double *eblk, double *vals, int mode,
                            hipStream_t st);

} // namespace rg
