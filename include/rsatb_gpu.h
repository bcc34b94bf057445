/*
 * rsatb_gpu.h -- C ABI of the MI355X (gfx950) stochastic Chebyshev/KPM engine.
 *
 * Plain C: no exceptions cross this boundary, every call returns an int
 * status (RG_OK = 0) and rg_last_error() holds the message of the last
 * failure on that context. One HIP stream per context; every call is
 * synchronous on return. Host arrays are borrowed for the duration of the
 * call only; device buffers are owned by the context.
 *
 * Reference seams replaced (paths relative to the RSATB source tree):
 *   - the operator:   Chebyshev<T> ctor + set_window + apply_op/apply_form
 *                     rsatb/calc/cheby.hpp:167-201, 228-235, 517-520, 555-591
 *   - the matrices:   BlockSpMat<T>::mult_dense / mult_dense_adjoint, reached
 *                     through Chebyshev::mult_by (cheby.hpp:599-606) and
 *                     ISQRTMPreconditioner::apply/apply_adjoint
 *                     (rsatb/calc/preconditioner.hpp:213-227)
 *   - the samples:    SampleVectors::initialize (rsatb/calc/observables.hpp:84-97)
 *                     and RademacherSign (rsatb/calc/rs_ldos.hpp:256-264)
 *   - the hot loop:   kpm_dos_moments multivector (rsatb/calc/dos.hpp:124-189)
 *   - the DOS task:   ComputeDos kpm branch (rsatb/calc/observables.hpp:311-444)
 *                     with evaluate_kpm_density / jackson_kernel
 *                     (rsatb/calc/kpm.hpp:18-26, 52-85, 206-241)
 *   - interruption:   util::throw_if_interrupted (rsatb/utils/interrupt.hpp:100-107)
 *
 * Host vector format of every vector argument: row-major [rows][R], either
 * real doubles or complex doubles interleaved (re, im), as the caller states.
 * Matrices: BSR with uniform b x b blocks, int64 block-row pointers, int32
 * block-column indices, entries row-major within each block, blocks in CSR
 * order -- the layout rsatb's dump_operator_coo walks
 * (rsatb/pybind_rsatb.cpp:636-668).
 */
#ifndef RSATB_GPU_H
#define RSATB_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#if defined(__GNUC__) || defined(__clang__)
#define RG_API __attribute__((visibility("default")))
#else
#define RG_API
#endif

typedef struct rg_context rg_context;

/* status codes */
#define RG_OK 0
#define RG_ERR_INVALID 1   /* bad argument or inconsistent shapes */
#define RG_ERR_HIP 2       /* a HIP runtime call failed */
#define RG_ERR_STATE 3     /* call out of order (operator or samples missing) */
#define RG_ERR_CANCELLED 4 /* rg_cancel() was observed between two launches */
#define RG_ERR_NOMEM 5     /* device allocation failed */
#define RG_ERR_UNSUPPORTED 6

/* matrix slots: H, Z (applied as Z x) and Z^H (applied as Z^H x), the three
 * factors of the sandwich Z^H H Z (cheby.hpp:563-568). Z^H is stored as its
 * own BSR instead of being reached through a scatter kernel. */
#define RG_SLOT_H 0
#define RG_SLOT_Z 1
#define RG_SLOT_ZH 2
#define RG_SLOT_S 3 /* overlap input of rg_isqrt_patches */
/* bond-moment velocity factors B_a (V_a = i B_a) of rg_sk_velocity, read by
 * rg_kubo_greenwood */
#define RG_SLOT_VX 4
#define RG_SLOT_VY 5
#define RG_SLOT_VZ 6
/* tangent direction of rg_dc_spectrum_tangent: dH and dB_a (real, H's
 * block rows and block size; any pattern) */
#define RG_SLOT_DH 7
#define RG_SLOT_DVX 8
#define RG_SLOT_DVY 9
#define RG_SLOT_DVZ 10
#define RG_SLOT_COUNT 11

/* value types of an uploaded matrix (FP32 / complex-FP32 values are widened
 * to FP64 exactly on upload) */
#define RG_F64 0
#define RG_F32 1
#define RG_C128 2
#define RG_C64 3

/* operator forms (ChebyOpForm, cheby.hpp:78-82; Solve is not provided) */
#define RG_FORM_ORTHOGONAL 0
#define RG_FORM_SANDWICH 1

/* sample kinds */
#define RG_SAMPLES_RADEMACHER 0     /* RademacherSign(seed,row,col)/sqrt(rows), real */
#define RG_SAMPLES_PHASE 1          /* e^{2 pi i u(seed,row,col)}/sqrt(rows), complex */
#define RG_SAMPLES_UPLOAD_REAL 2    /* host_ptr: [rows][R] doubles */
#define RG_SAMPLES_UPLOAD_COMPLEX 3 /* host_ptr: [rows][R][2] doubles */

/* arithmetic of the moment loop */
#define RG_PREC_FP64 0
#define RG_PREC_FP32 1 /* FP32 SpMM + recurrence; moment dots of exact FP32 products in FP64,
                                compensated (Neumaier) across workgroup partials */

/* doubling: product-identity two moments per apply (dos.hpp:150-173) */
#define RG_DOUBLING_AUTO -1 /* on: both provided forms are Hermitian (cheby.hpp:222) */
#define RG_DOUBLING_OFF 0
#define RG_DOUBLING_ON 1

/* KPM kernels (KpmKernel, kpm.hpp:49) */
#define RG_KERNEL_NONE 0
#define RG_KERNEL_JACKSON 1
#define RG_KERNEL_COSTAPER 2
#define RG_KERNEL_VPTAPER 3

/* ---- context (no reference counterpart: the device handle) ---- */
RG_API int rg_create(int device_id, rg_context **ctx);
RG_API void rg_destroy(rg_context *ctx);
RG_API const char *rg_last_error(const rg_context *ctx);

/* Free and total device memory of the context's device after its stream has
 * drained (hipMemGetInfo): for callers that size their problems or watch the
 * device memory a long fitting loop holds. No counterpart in the reference
 * (a host-memory code). */
RG_API int rg_mem_info(rg_context *ctx, size_t *free_bytes, size_t *total_bytes);
/* Runtime version of this library, for the loader's sanity check. */
RG_API int rg_version(void);

/* ---- operands ---- */

/* Replaces BlockSpMat<T> assembly for one operator slot (H, Z, Z^H, S):
 * rsatb/setup/operator_set.hpp:291-298 produces H(k)/S(k), and
 * preconditioner.hpp:184-196 owns Z. `values` holds nnzb*b*b doubles (F64),
 * floats (F32), or nnzb*b*b (re, im) pairs of doubles (C128) or floats (C64),
 * each block row-major. */
RG_API int rg_upload_bsr(rg_context *ctx, int slot, int64_t n_block_rows, int block_dim,
                  const int64_t *row_ptr, const int32_t *col_ind, const void *values,
                  int dtype);

/* Chebyshev ctor + set_window (cheby.hpp:167-186, 517-520): the recurrence
 * iterates (A - center)/half_width with A = H or Z^H H Z. */
RG_API int rg_set_operator(rg_context *ctx, int form, double center, double half_width);

/* SampleVectors::initialize (observables.hpp:84-97). host_ptr is read only
 * for the UPLOAD kinds. */
RG_API int rg_set_samples(rg_context *ctx, int64_t n_rows, int R, int kind, uint64_t seed,
                   const double *host_ptr);

/* Copies the current sample block to the host, [rows][R] real or
 * [rows][R][2] complex as `as_complex` asks (checks bit-exactness). */
RG_API int rg_get_samples(rg_context *ctx, double *host_out, int as_complex);

/* ---- operator application (Chebyshev::apply_op, cheby.hpp:228-235) ---- */

/* y = (A x - center x)/half_width for R columns of host vectors. */
RG_API int rg_apply_op(rg_context *ctx, int R, const double *x, double *y, int is_complex);

/* y = M x for one stored slot (BlockSpMat::mult_dense, cheby.hpp:599-606). */
RG_API int rg_bsr_mult(rg_context *ctx, int slot, int R, const double *x, double *y, int is_complex);

/* ---- the hot loop (kpm_dos_moments multivector, dos.hpp:124-189) ---- */

/* mu_out: [M] complex (re, im), the mean over the R sample columns.
 * mu_cols_out: optional [M][R] complex per-column moments (NULL to skip). */
RG_API int rg_kpm_moments(rg_context *ctx, int M, int precision, int doubling, double *mu_out,
                   double *mu_cols_out);

/* No counterpart in the reference. rg_kpm_moments for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): the same kernels and the same results
 * (bit for bit), with the work vectors, dot partials and per-step moment store
 * in a fixed array of 128 MiB that the library keeps on every device it runs
 * on until the process ends; the call makes no hipMalloc/hipFree. The caller
 * still provides the inputs (matrices, window, samples: set beforehand with
 * the allocating calls) and mu_out / mu_cols_out. RG_ERR_UNSUPPORTED when the
 * workspace ((3 + 2 sandwich + 1 plain) vectors of rows*R values, 2048*3*R
 * complex partials and (M/2 + 1)*3*R complex dots, 256-byte aligned) exceeds
 * the array; RG_ERR_STATE for RG_PREC_FP32 before an FP32 rg_kpm_moments made
 * the FP32 matrix copies. No forward states are kept (the backward needs
 * rg_kpm_moments) and no timing events are recorded. All contexts on one
 * device share the array: calls that overlap on one device race. */
RG_API int rg_kpm_moments_noalloc(rg_context *ctx, int M, int precision, int doubling, double *mu_out,
                                  double *mu_cols_out);

/* ---- projected moments (kpm_pdos_moments multivector, pdos.hpp:149-266) ---- */

/* The moments of rg_kpm_moments resolved onto groups of rows from the same
 * recurrence: mu_m^g = mean over columns of <t|P_g|t'>, P_g the diagonal
 * projector on the rows r with group_of_row[r] == g (a negative id puts a row
 * in no group; ids must be < n_groups, pdos.hpp:159-167). n_rows must equal
 * the sample rows. Doubling as rg_kpm_moments, applied per group (the masked
 * identity of pdos.hpp:29-36).
 * total_out: [M] complex, the all-rows moments (= rg_kpm_moments' mu_out).
 * group_out: [n_groups][M] complex, row-major (PdosMoments::group layout,
 * pdos.hpp:129-141). With a partition of the rows, sum_g group = total to
 * rounding. */
RG_API int rg_kpm_pdos_moments(rg_context *ctx, int M, int precision, int doubling, int64_t n_rows,
                        const int32_t *group_of_row, int n_groups, double *total_out,
                        double *group_out);

/* ---- adjoint (no reference counterpart: the differentiable form) ---- */

/* Keeps the forward states of the next doubled FP64 rg_kpm_moments call for
 * rg_kpm_moments_backward: psi_0, psi_1 and every pair (psi_{kK-1}, psi_{kK}),
 * K = stride (stride 1 keeps every state; 0 turns keeping off). */
RG_API int rg_set_checkpointing(rg_context *ctx, int stride);

/* Reverse mode through the last kept moment run, for a real loss L of the
 * column-mean moments: dL_dmu[m] = dL/dRe(mu_m), m < M. Writes dL/dH for every
 * stored entry of the real H in RG_SLOT_H (nnzb*b*b doubles, row-major within
 * blocks), the window (center, half_width) held fixed; for the sandwich form
 * Z is held fixed (rg_kpm_moments_backward_ex also differentiates Z and S).
 * RG_ERR_STATE when the operator, the window or the samples changed after the
 * forward pass. The adjoint states obey
 *   lam_n = 2 A~ lam_{n+1} - lam_{n+2} + dL/dpsi_n (direct),
 * dL/dA~ = lam_1 psi_0^T + 2 sum_{n>=2} lam_n psi_{n-1}^T, summed over columns;
 * segments between checkpoints are recomputed forward on the device. */
RG_API int rg_kpm_moments_backward(rg_context *ctx, const double *dL_dmu, double *grad_out);

/* The same reverse pass with the gradient per BSR slot (the per-slot
 * dL_dvalues of the differentiable form); each output is optional (null
 * skips it), at least one is required:
 *   grad_h: dL/dH on RG_SLOT_H's pattern, as rg_kpm_moments_backward;
 *   grad_z: dL/dZ on RG_SLOT_Z's pattern (sandwich form, real Z), with the
 *           two factors tied (Z^H = Z^T): dL/dZ = (H Z G + H Z G^T) / a,
 *           G = dL/dA~ as above;
 *   grad_s: dL/dS on RG_SLOT_S's pattern through the graph-local factor
 *           (graph_inverse_sqrt's seed, rsatb/calc/isqrtm.hpp:149-232, as
 *           rg_isqrt_patches builds it): needs Z and Z^H from rg_isqrt_patches
 *           of the overlap now in RG_SLOT_S. Per patch A_i = S[C_i, C_i],
 *           dL/dA_i = L_f(A_i, E_i) with f(x) = x^{-1/2} (the Frechet
 *           derivative, self-adjoint for symmetric A_i) and E_i = dL/dZ's
 *           block row i in the rows of block i, from the tangent of the same
 *           batched Newton-Schulz iteration on the device; every patch entry is
 *           added to the stored block it came from.
 * All outputs nnzb*b*b doubles, row-major within blocks; per-entry values
 * whose symmetric part (G_ij + G_ji) is what a symmetric parameterization
 * sees. RG_ERR_STATE when H, Z, Z^H, S, the window or the samples changed
 * after the forward pass. */
RG_API int rg_kpm_moments_backward_ex(rg_context *ctx, const double *dL_dmu, double *grad_h, double *grad_z,
                                      double *grad_s);

/* ---- DOS task (ComputeDos kpm branch, observables.hpp:311-444) ---- */

/* Grid and moment count ComputeDos derives from (emin, emax, dE):
 * nt = ceil(span/dE), N = ceil(sqrt(8 ln 2) pi half_width / dE)
 * (observables.hpp:336-350); the output has 2*nt points. */
RG_API int rg_dos_plan(double emin, double emax, double de, int *n_points, int *n_moments);

/* Runs the moment loop and the Jackson reconstruction; dos_out and
 * omega_out hold 2*nt values (rg_dos_plan). dos = 2 rho (observables.hpp:388-393). */
RG_API int rg_compute_dos(rg_context *ctx, double emin, double emax, double de, int precision,
                   double *dos_out, double *omega_out);

/* evaluate_kpm_density (kpm.hpp:206-241): host FP64. mu: [M] complex. */
RG_API int rg_kpm_density(const double *mu, int M, const double *x, int64_t nx, double edge_inset,
                   int kernel, double *rho_out);

/* The MFMA SpMM plan the library derives for a square block pattern (real
 * slots with b >= 8 get one at every write): sched_out[nbr] the block rows
 * in processing order (regions of *sreg_out rows, one XCD's rows in flight,
 * each grown breadth-first over the block graph), kperm_out[nnzb] the stored
 * block every position of a row takes (blocks ordered by a distance-2
 * colouring of the block columns). Host only, needs no device. No
 * counterpart in the reference (an execution order internal to this
 * library); for inspection and tests. */
RG_API int rg_mfma_plan(int64_t n_block_rows, const int64_t *row_ptr, const int32_t *col_ind,
                        int32_t *sched_out, int32_t *kperm_out, int32_t *sreg_out);

/* ---- operand inspection (dump_operator_coo, rsatb/pybind_rsatb.cpp:612-670) ---- */

RG_API int rg_bsr_info(rg_context *ctx, int slot, int64_t *n_block_rows, int *block_dim,
                       int64_t *nnzb, int *is_complex);
/* Copies a slot back in the upload layout (values row-major within blocks);
 * any output pointer may be NULL. */
RG_API int rg_download_bsr(rg_context *ctx, int slot, int64_t *row_ptr, int32_t *col_ind,
                           double *values);

/* ---- Slater-Koster assembly on the device (SkTableConstruction::build_images,
 *      rsatb/model/sktb/sktb_construction.hpp:327-425, over from_graph arrays,
 *      rsatb/pybind_rsatb.cpp:562-593) ---- */

/* One matrix (H or S) of a resolved SK table (sk_table.hpp:104-161), as
 * plain arrays. Every type has norb orbitals (the device BSR is uniform).
 * Orbital o of type t: angular momentum orb_l[t norb + o] (0, 1, 2) and
 * magnetic index orb_m[t norb + o] (DeePTB order: py pz px; dxy dyz dz2 dzx
 * dx2-y2); onsite[t norb + o] its onsite diagonal. pair_cutoff[rt T + ct]:
 * the largest cutoff of this matrix's channels of the ordered type pair, 0
 * when it has none. Channel c: types, shells (row_l, col_l), bond (0 sigma,
 * 1 pi, 2 delta, at most min(row_l, col_l)), radial kind (0 natural cubic
 * spline through rad_x/rad_y with second derivatives rad_y2; 1 exp_poly with
 * coefficients rad_y, decay lambda and cosine taper from r_on), its data at
 * rad_*[ch_off .. ch_off + ch_n), and cutoff ch_r_cut. A shell pair l < l'
 * appears as both (l, l') and (l', l) (sk_table.hpp:404-407). */
typedef struct rg_sk_table {
    int32_t n_types, norb, n_channels;
    const int32_t *orb_l, *orb_m;
    const double *onsite;
    const double *pair_cutoff;
    const int32_t *ch_row_type, *ch_col_type, *ch_row_l, *ch_col_l, *ch_bond, *ch_kind, *ch_off, *ch_n;
    const double *ch_r_cut, *ch_lambda, *ch_r_on;
    const double *rad_x, *rad_y, *rad_y2;
    int64_t n_rad;
} rg_sk_table;

/* Builds H(k) or S(k) of the table into `slot` on the device: the onsite
 * diagonal block of every atom, and for each edge (edge_index[2e] ->
 * edge_index[2e + 1], bond vector pos[j] + shift . cell - pos[i], cell rows
 * the lattice vectors) the SK block of the table pair of its types, else the
 * transposed block of the reversed pair (sktb_construction.hpp:476-500);
 * edges whose block is empty or has no entry above 1e-14 are not stored
 * (:385). Blocks of the same (i, j) are summed with phases exp(2 pi i k.shift)
 * (kpoint NULL or zero: real values). The slot then serves the moment loop
 * like an uploaded matrix; *nnzb_out receives the stored block count. */
RG_API int rg_sk_assemble(rg_context *ctx, int slot, const rg_sk_table *table, int64_t n_atoms,
                          const int32_t *atom_type, const double *pos, const double *cell,
                          int64_t n_edges, const int32_t *edge_index, const int32_t *edge_shift,
                          const double *kpoint, int64_t *nnzb_out);

/* ---- Kubo-Greenwood DC conductivity (dc_spectrum_fftcheb, Greenwood formulation,
 *      rsatb/calc/fftcheb_transport.hpp:543-623; velocity of an orthogonal basis,
 *      rsatb/operator/commutator_velocity.hpp:7-43) ---- */

/* The commutator velocity V_a = i B_a of an orthogonal SK model:
 * B_a = sum over edges of (pos_j + shift . cell - pos_i)_a times the edge's
 * SK block (AccumulateBondMoment with the displacement weight and factor i),
 * into RG_SLOT_VX .. RG_SLOT_VZ on H's stored pattern. At kpoint (fractional,
 * NULL or zero: Gamma, real factors) each edge also carries the Bloch phase
 * e^{2 pi i k . shift} that rg_sk_assemble gives H(k), so B_a(k) is complex
 * anti-Hermitian and V_a(k) the velocity in H(k)'s gauge (the reference's
 * atomic-gauge V and H(k), commutator_velocity.hpp:24-31, are the same pair
 * conjugated by diag(e^{i k . r})). Same arguments as rg_sk_assemble with the
 * Hamiltonian table. */
RG_API int rg_sk_velocity(rg_context *ctx, const rg_sk_table *table, int64_t n_atoms,
                          const int32_t *atom_type, const double *pos, const double *cell,
                          int64_t n_edges, const int32_t *edge_index, const int32_t *edge_shift,
                          const double *kpoint);

/* Greenwood conductivity spectrum of the orthogonal form (H in RG_SLOT_H,
 * velocity factors in RG_SLOT_VX.., window of rg_set_operator, samples of
 * rg_set_samples) with M moments and damping `kernel` (RG_KERNEL_*):
 * sigma_out[k * 3 + a] = prefactor * kernel_aa(eps) on the ascending
 * Chebyshev nodes, energy_out[k] their energies (may be NULL). The left
 * and right Chebyshev sequences of every sample are kept on the device and
 * contracted as mu^a = L^H R (one GEMM per axis); mu_out (may be NULL)
 * receives [3][M][M] complex moments scaled by 1 / (2 a)^2 and summed over
 * samples. Memory: 6 M x rows x R vector elements. */
RG_API int rg_kubo_greenwood(rg_context *ctx, int M, int kernel, double prefactor, double *sigma_out,
                             double *energy_out, double *mu_out);

/* The full DC spectrum of dc_spectrum_fftcheb (fftcheb_transport.hpp:543-653)
 * for the same inputs: formulation 0 Greenwood (the diagonal components) or
 * 1 Bastin (all six, integrated over the energy: sigma at T = 0 as a function
 * of the Fermi energy); sigma_out[k * 6 + comp], comp in (xx, xy, xz, yy, yz,
 * zz); mu_out (may be NULL): [pairs][M][M] complex, 3 pairs (xx, yy, zz) for
 * Greenwood, 6 in component order for Bastin. */
RG_API int rg_dc_spectrum(rg_context *ctx, int M, int formulation, int kernel, double prefactor,
                          double *sigma_out, double *energy_out, double *mu_out);

/* rg_dc_spectrum with a precision (RG_PREC_FP64 or RG_PREC_FP32). FP32: the
 * recurrence and the velocity products run on FP32 vectors and matrices and
 * the stored sequences are FP32 (half the memory); the moment GEMMs widen
 * K-chunks of them to FP64 and accumulate in FP64. The reference's transport
 * path is FP64 only (dc_spectrum_fftcheb, fftcheb_transport.hpp:543-653);
 * this is the precision form of that spectrum. */
RG_API int rg_dc_spectrum_ex(rg_context *ctx, int M, int formulation, int kernel, double prefactor, int precision,
                             double *sigma_out, double *energy_out, double *mu_out);

/* rg_dc_spectrum and its directional derivative: dsigma_out[k * 6 + comp] =
 * d sigma / d e along H -> H + e dH, B_a -> B_a + e dB_a (RG_SLOT_DH,
 * RG_SLOT_DVX..), with the window of rg_set_operator held fixed. The
 * tangent recurrence yd_n = 2 (dH / a) y_{n-1} + 2 A yd_{n-1} - yd_{n-2} runs
 * beside each stored sequence, and dmu = s^2 (dL^H R + L^H dR). For a
 * parameterization H(theta) = sum_k theta_k H_k (and B_a likewise) one call
 * per parameter gives the Jacobian d sigma / d theta, the derivative behind
 * gradients of conductivity losses. No counterpart in the reference (its
 * transport path has no derivatives); the spectrum it differentiates is
 * dc_spectrum_fftcheb (fftcheb_transport.hpp:543-653). Gamma only (real H and
 * tangent matrices). Memory: 12 M x rows x R vector elements. */
RG_API int rg_dc_spectrum_tangent(rg_context *ctx, int M, int formulation, int kernel, double prefactor,
                                  double *sigma_out, double *dsigma_out, double *energy_out);

/* ---- graph-local S^{-1/2} (graph_inverse_sqrt seed, rsatb/calc/isqrtm.hpp:137-232;
 *      ISQRTMPreconditioner ctor, rsatb/calc/preconditioner.hpp:184-192) ---- */

/* From the overlap in RG_SLOT_S (real, symmetric pattern, diagonal blocks
 * stored) builds Z in RG_SLOT_Z and Z^H in RG_SLOT_ZH: block row i of Z is
 * block i's rows of S[C_i, C_i]^{-1/2}, C_i the stored columns of row i, by
 * batched FP64 Newton-Schulz on the device. Converged when
 * ||I - Z Y||_F < tol on every patch; fails (RG_ERR_INVALID) when an
 * iteration diverges (residual > 100, isqrtm.hpp:637-643) or maxit is hit. */
RG_API int rg_isqrt_patches(rg_context *ctx, double tol, int maxit, int *iters_out,
                            double *resid_out);

/* ---- spectral window (EstimateSpectralBounds, rsatb/setup/spectral_bounds.hpp:445-489;
 *      the extremes come from FP64 block Lanczos on the device instead of PRIMME) ---- */

/* 16 independent Lanczos chains of `steps` steps from counter-based start
 * vectors (seed); returns [min Ritz - padding, max Ritz + padding]. Needs the
 * factors of `form` uploaded; the window set by rg_set_operator is not used. */
RG_API int rg_estimate_bounds(rg_context *ctx, int form, int steps, uint64_t seed,
                              double padding, double *emin, double *emax);

/* ---- interruption (throw_if_interrupted, interrupt.hpp:100-107) ---- */

/* Sets a flag that the moment loop checks between launches; the running
 * call then returns RG_ERR_CANCELLED. Safe from another thread. The flag is
 * cleared when the next loop starts. */
RG_API int rg_cancel(rg_context *ctx);

/* ---- measurement (no reference counterpart; for the benchmark) ---- */

/* Turns per-launch HIP event timing on (1) or off (0) and clears stats. */
RG_API int rg_set_profiling(rg_context *ctx, int on);

/* Kernel classes timed when profiling is on. */
#define RG_KCLASS_SPMM_PLAIN 0 /* BSR SpMM, Y = M X (sandwich inner factors) */
#define RG_KCLASS_SPMM_STEP 1  /* SpMM fused with the recurrence and dot partials */
#define RG_KCLASS_REDUCE 2     /* dot-partial reduction */
#define RG_KCLASS_GROUP 3      /* group-resolved dots of the projected moments */
#define RG_KCLASS_KG_GEMM 4    /* Kubo-Greenwood moment GEMM mu = L^H R (one axis) */
#define RG_KCLASS_COUNT 5

/* Average launch duration (ms), launch count, and the algorithmic bytes and
 * flops of one launch of that class as last launched. */
RG_API int rg_kernel_stats(rg_context *ctx, int kclass, double *avg_ms, int64_t *count,
                    double *bytes_per_launch, double *flops_per_launch);

#ifdef __cplusplus
}
#endif

#endif /* RSATB_GPU_H */
