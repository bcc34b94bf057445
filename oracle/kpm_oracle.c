/*
 * kpm_oracle.c -- CPU FP64 restatement of RSATB's stochastic KPM/Chebyshev
 * path. TEST INFRASTRUCTURE ONLY: this file is the checker that the HIP
 * library is compared against, and the CPU baseline timed beside it. Only
 * tests, the smoke check and the benchmark's cpu_baseline leg load it; the
 * product (rsatb_amd) never links, imports or calls it.
 *
 * Everything is stored the way the reference stores it: std::complex<double>
 * for matrix entries and vectors (rsatb/rsatb.hpp:53-57,
 * rsatb/calc/observables.hpp:51), interleaved (re, im) here.
 *
 * Layout contracts (the vbcsr BSR contract as far as rsatb pins it):
 *   - BSR with uniform b x b blocks, CSR over block rows, entries row-major
 *     within a block (rsatb/calc/test/test_preconditioner.cpp:1-10,37-80;
 *     dump_operator_coo indexing rsatb/pybind_rsatb.cpp:654-667).
 *   - multivectors row-major [rows][R] (rsatb/calc/dos.hpp:120-123).
 *   - dot(a, b)[c] = sum_i conj(a[i,c]) * b[i,c] (rsatb/calc/pdos.hpp:145-161).
 *
 * Functions and the reference lines they restate:
 *   or_mix64, or_rademacher_sign ....... rsatb/calc/rs_ldos.hpp:249-264
 *   or_bsr_mult ........................ vbcsr BlockSpMat::mult_dense, row gather
 *                                        (called at rsatb/calc/cheby.hpp:599-606)
 *   or_bsr_mult_adjoint ................ vbcsr mult_dense_adjoint, scatter
 *                                        (rsatb/calc/preconditioner.hpp:225-227)
 *   or_apply_op ........................ Chebyshev::apply_form, Orthogonal and
 *                                        Sandwich (rsatb/calc/cheby.hpp:555-591)
 *   or_kpm_moments (fused = 0) ......... kpm_dos_moments multivector
 *                                        (rsatb/calc/dos.hpp:124-189), the
 *                                        doubled and the plain recurrences
 *   or_kpm_pdos_moments ................ kpm_pdos_moments multivector
 *                                        (rsatb/calc/pdos.hpp:149-266)
 *   or_jackson_kernel .................. rsatb/calc/kpm.hpp:18-26
 *   or_kpm_kernel_weights .............. rsatb/calc/kpm.hpp:52-85
 *   or_kpm_density ..................... rsatb/calc/kpm.hpp:206-241
 *
 * or_kpm_moments(fused = 1) is the same arithmetic with the scale, shift,
 * recurrence and dots folded into the last operator pass: the "fused CPU"
 * baseline, not a separate algorithm.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
    int64_t nbr;            /* block rows (= block columns: square) */
    int32_t b;              /* uniform block dimension */
    const int64_t *row_ptr; /* [nbr + 1] */
    const int32_t *col_ind; /* [nnzb] */
    const double *vals;     /* [nnzb][b][b][2], complex interleaved, row-major */
} or_bsr;

enum { OR_FORM_ORTHOGONAL = 0, OR_FORM_SANDWICH = 1 };

typedef struct {
    int32_t form;
    const or_bsr *H;
    const or_bsr *Z; /* sandwich: y = Z^H H Z x; Z^H applied by the scatter kernel */
    double center;
    double half_width;
    const or_bsr *ZH; /* optional explicit Z^H: row-gather instead of the scatter */
} or_op;

/* ---------------------------------------------------------------- RNG --- */

/* SplitMix64 finaliser (Steele, Lea & Flood 2014), rs_ldos.hpp:256-261. */
uint64_t or_mix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

/* rs_ldos.hpp:256-264: sign of mix(mix(mix(seed) ^ row) ^ state). */
double or_rademacher_sign(uint64_t seed, uint64_t row, uint64_t state) {
    const uint64_t h = or_mix64(or_mix64(or_mix64(seed) ^ row) ^ state);
    return (h & 1ull) ? 1.0 : -1.0;
}

/* R Rademacher columns, RademacherSign(seed, row, col) / sqrt(nrows). */
void or_samples_rademacher(uint64_t seed, int64_t nrows, int32_t R, double *out) {
    const double s = 1.0 / sqrt((double)nrows);
    for (int64_t i = 0; i < nrows; ++i)
        for (int32_t c = 0; c < R; ++c) {
            out[2 * (i * R + c)] = or_rademacher_sign(seed, (uint64_t)i, (uint64_t)c) * s;
            out[2 * (i * R + c) + 1] = 0.0;
        }
}

/* Unit-norm random phase e^{i phi}/sqrt(nrows) (observables.hpp:54-97 draws
 * the same distribution through vbcsr set_random, whose bit stream is not in
 * the tree). phi = 2 pi u, u = top 53 bits of the RademacherSign hash chain
 * times 2^-53: this build's own definition of the counter-based phase draw. */
void or_samples_phase(uint64_t seed, int64_t nrows, int32_t R, double *out) {
    const double s = 1.0 / sqrt((double)nrows);
    const double two_pi = 6.283185307179586476925286766559;
    for (int64_t i = 0; i < nrows; ++i)
        for (int32_t c = 0; c < R; ++c) {
            const uint64_t h = or_mix64(or_mix64(or_mix64(seed) ^ (uint64_t)i) ^ (uint64_t)c);
            const double u = (double)(h >> 11) * (1.0 / 9007199254740992.0);
            out[2 * (i * R + c)] = cos(two_pi * u) * s;
            out[2 * (i * R + c) + 1] = sin(two_pi * u) * s;
        }
}

/* --------------------------------------------------------- BSR kernels --- */

/* Y = A X, row gather. X, Y: [nbr*b][R] complex. */
void or_bsr_mult(const or_bsr *A, const double *X, int32_t R, double *Y) {
    const int32_t b = A->b;
    const int64_t nbr = A->nbr;
#pragma omp parallel for schedule(static)
    for (int64_t br = 0; br < nbr; ++br) {
        double *yrow = Y + 2 * (br * b) * (int64_t)R;
        memset(yrow, 0, sizeof(double) * 2 * (size_t)b * (size_t)R);
        for (int64_t s = A->row_ptr[br]; s < A->row_ptr[br + 1]; ++s) {
            const int64_t bc = A->col_ind[s];
            const double *blk = A->vals + 2 * s * (int64_t)b * b;
            const double *xcol = X + 2 * (bc * b) * (int64_t)R;
            for (int32_t i = 0; i < b; ++i) {
                double *yi = yrow + 2 * (int64_t)i * R;
                for (int32_t j = 0; j < b; ++j) {
                    const double ar = blk[2 * (i * b + j)], ai = blk[2 * (i * b + j) + 1];
                    const double *xj = xcol + 2 * (int64_t)j * R;
                    for (int32_t c = 0; c < R; ++c) {
                        const double xr = xj[2 * c], xi = xj[2 * c + 1];
                        yi[2 * c] += ar * xr - ai * xi;
                        yi[2 * c + 1] += ar * xi + ai * xr;
                    }
                }
            }
        }
    }
}

/* Y = A^H X by scattering each stored block's conjugate transpose (the
 * reference's mult_dense_adjoint route). Serial: the scatter is the kernel
 * the reference measured as not thread-scalable (cheby.hpp:38-56). */
void or_bsr_mult_adjoint(const or_bsr *A, const double *X, int32_t R, double *Y) {
    const int32_t b = A->b;
    const int64_t nbr = A->nbr;
    memset(Y, 0, sizeof(double) * 2 * (size_t)(nbr * b) * (size_t)R);
    for (int64_t br = 0; br < nbr; ++br) {
        const double *xrow = X + 2 * (br * b) * (int64_t)R;
        for (int64_t s = A->row_ptr[br]; s < A->row_ptr[br + 1]; ++s) {
            const int64_t bc = A->col_ind[s];
            const double *blk = A->vals + 2 * s * (int64_t)b * b;
            double *ycol = Y + 2 * (bc * b) * (int64_t)R;
            /* (A^H)[bc*b + j][br*b + i] = conj(A[br*b + i][bc*b + j]) */
            for (int32_t j = 0; j < b; ++j) {
                double *yj = ycol + 2 * (int64_t)j * R;
                for (int32_t i = 0; i < b; ++i) {
                    const double ar = blk[2 * (i * b + j)], ai = -blk[2 * (i * b + j) + 1];
                    const double *xi = xrow + 2 * (int64_t)i * R;
                    for (int32_t c = 0; c < R; ++c) {
                        const double xr = xi[2 * c], xim = xi[2 * c + 1];
                        yj[2 * c] += ar * xr - ai * xim;
                        yj[2 * c + 1] += ar * xim + ai * xr;
                    }
                }
            }
        }
    }
}

/* ---------------------------------------------------- vector passes --- */

static void v_axpy(int64_t n, double a, const double *x, double *y) { /* y += a x */
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < 2 * n; ++k) y[k] += a * x[k];
}
static void v_scale(int64_t n, double a, double *y) {
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < 2 * n; ++k) y[k] *= a;
}
static void v_axpby(int64_t n, double a, const double *x, double b, double *y) { /* y = a x + b y */
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < 2 * n; ++k) y[k] = a * x[k] + b * y[k];
}

/* dots[c] = sum_i conj(a[i,c]) b[i,c]. Each thread sums a static block of
 * rows into its own slot; the slots are then added in thread order, so the
 * result is the same on every run with the same thread count. */
static void v_dot(int64_t rows, int32_t R, const double *a, const double *b, double *dots) {
    const int nt = omp_get_max_threads();
    double *part = (double *)calloc((size_t)nt * (size_t)(2 * R), sizeof(double));
#pragma omp parallel num_threads(nt)
    {
        double *p = part + (size_t)omp_get_thread_num() * (size_t)(2 * R);
#pragma omp for schedule(static)
        for (int64_t i = 0; i < rows; ++i) {
            const double *ar = a + 2 * i * R, *brw = b + 2 * i * R;
            for (int32_t c = 0; c < R; ++c) {
                const double xr = ar[2 * c], xi = ar[2 * c + 1];
                const double yr = brw[2 * c], yi = brw[2 * c + 1];
                p[2 * c] += xr * yr + xi * yi;
                p[2 * c + 1] += xr * yi - xi * yr;
            }
        }
    }
    for (int32_t c = 0; c < 2 * R; ++c) dots[c] = 0.0;
    for (int t = 0; t < nt; ++t)
        for (int32_t c = 0; c < 2 * R; ++c) dots[c] += part[(size_t)t * (size_t)(2 * R) + c];
    free(part);
}

/* ------------------------------------------------------- apply_form --- */

/* y = (A x - center x) / half_width; A = H or Z^H H Z (cheby.hpp:555-591).
 * tmp: scratch of the same shape (cheby.hpp:616-628). */
int or_apply_op(const or_op *op, const double *x, int32_t R, double *y, double *tmp) {
    const int64_t n = op->H->nbr * op->H->b * (int64_t)R;
    if (op->form == OR_FORM_ORTHOGONAL) {
        or_bsr_mult(op->H, x, R, y);
    } else if (op->form == OR_FORM_SANDWICH) {
        or_bsr_mult(op->Z, x, R, y);         /* M->apply(x, y) */
        or_bsr_mult(op->H, y, R, tmp);       /* mult_by(H, y, tmp) */
        if (op->ZH) or_bsr_mult(op->ZH, tmp, R, y);
        else or_bsr_mult_adjoint(op->Z, tmp, R, y); /* M->apply_adjoint(tmp, y) */
    } else {
        return -1;
    }
    v_axpy(n, -op->center, x, y);
    v_scale(n, 1.0 / op->half_width, y);
    return 0;
}

/* ------------------------------------------------------------- moments --- */

static void col_mean(int32_t R, const double *dots, double *out) {
    double sr = 0.0, si = 0.0;
    for (int32_t c = 0; c < R; ++c) { sr += dots[2 * c]; si += dots[2 * c + 1]; }
    out[0] = sr / (double)R;
    out[1] = si / (double)R;
}

/* Fused last pass for the CPU baseline: t2 = (2/a)(A t1 - b t1) - t0 and the
 * dots <t2|t2>, <t2|t1> in the same sweep over rows. av = A t1 precomputed
 * (orthogonal: computed here row by row). */
static void fused_step(const or_op *op, const double *t0, const double *t1, double *t2,
                       const double *av, int32_t R, double *d22, double *d21) {
    const int64_t rows = op->H->nbr * op->H->b;
    const double alpha = 2.0 / op->half_width, beta = -2.0 * op->center / op->half_width;
    const int nt = omp_get_max_threads();
    double *part = (double *)calloc((size_t)nt * (size_t)(4 * R), sizeof(double));
#pragma omp parallel num_threads(nt)
    {
        double *p22 = part + (size_t)omp_get_thread_num() * (size_t)(4 * R);
        double *p21 = p22 + 2 * R;
#pragma omp for schedule(static)
        for (int64_t i = 0; i < rows; ++i) {
            for (int32_t c = 0; c < R; ++c) {
                const int64_t k = 2 * (i * R + c);
                const double vr = alpha * av[k] + beta * t1[k] - t0[k];
                const double vi = alpha * av[k + 1] + beta * t1[k + 1] - t0[k + 1];
                t2[k] = vr;
                t2[k + 1] = vi;
                p22[2 * c] += vr * vr + vi * vi;
                p21[2 * c] += vr * t1[k] + vi * t1[k + 1];
                p21[2 * c + 1] += vr * t1[k + 1] - vi * t1[k];
            }
        }
    }
    /* thread slots added in thread order: the same result on every run */
    for (int32_t c = 0; c < 2 * R; ++c) { d22[c] = 0.0; d21[c] = 0.0; }
    for (int t = 0; t < nt; ++t) {
        const double *p22 = part + (size_t)t * (size_t)(4 * R), *p21 = p22 + 2 * R;
        for (int32_t c = 0; c < 2 * R; ++c) { d22[c] += p22[c]; d21[c] += p21[c]; }
    }
    free(part);
}

/* kpm_dos_moments, multivector (dos.hpp:124-189).
 *   psi: [rows][R] complex. mu: [M] complex (column mean).
 *   mu_cols: optional [M][R] complex per-column moments.
 *   doubling: 1 = Hermitian doubling (dos.hpp:153-173), 0 = plain (:174-188).
 *   fused: 0 = the reference's unfused passes, 1 = fused last pass.
 * Returns 0, or -1 on bad arguments. */
int or_kpm_moments(const or_op *op, const double *psi, int32_t R, int32_t M, int32_t doubling,
                   int32_t fused, double *mu, double *mu_cols) {
    if (M < 1 || R < 1) return -1;
    const int64_t rows = op->H->nbr * op->H->b;
    const int64_t n = rows * (int64_t)R;
    double *t0 = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    double *t1 = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    double *t2 = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    double *tmp = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    double *dots = (double *)malloc(sizeof(double) * 2 * (size_t)R);
    double *dots2 = (double *)malloc(sizeof(double) * 2 * (size_t)R);
    double *c0 = (double *)calloc(2 * (size_t)R, sizeof(double)); /* per-column mu_0 */
    double *c1 = (double *)calloc(2 * (size_t)R, sizeof(double)); /* per-column mu_1 */
    if (!t0 || !t1 || !t2 || !tmp || !dots || !dots2 || !c0 || !c1) return -1;
    memset(mu, 0, sizeof(double) * 2 * (size_t)M);
    if (mu_cols) memset(mu_cols, 0, sizeof(double) * 2 * (size_t)M * (size_t)R);

#define STORE_COLS(idx, src)                                                        \
    do {                                                                            \
        if (mu_cols) memcpy(mu_cols + 2 * (size_t)(idx) * R, (src), sizeof(double) * 2 * (size_t)R); \
    } while (0)

    memcpy(t0, psi, sizeof(double) * 2 * (size_t)n);
    v_dot(rows, R, t0, t0, dots); /* mu_0 = <t0|t0> (dos.hpp:141-143) */
    col_mean(R, dots, mu);
    memcpy(c0, dots, sizeof(double) * 2 * (size_t)R);
    STORE_COLS(0, dots);
    if (M > 1) {
        or_apply_op(op, t0, R, t1, tmp); /* t1 = A~ t0; mu_1 = <t0|t1> (:146-148) */
        v_dot(rows, R, t0, t1, dots);
        col_mean(R, dots, mu + 2);
        memcpy(c1, dots, sizeof(double) * 2 * (size_t)R);
        STORE_COLS(1, dots);
    }
    if (M > 1 && doubling) {
        /* t1 = psi_n, t0 = psi_{n-1} at the top of each pass (:153-173) */
        for (int n_ = 1;; ++n_) {
            if (!(fused && n_ > 1)) {
                if (2 * n_ - 1 < M && 2 * n_ - 1 > 1) {
                    v_dot(rows, R, t1, t0, dots);
                    for (int32_t c = 0; c < R; ++c) {
                        dots[2 * c] = 2.0 * dots[2 * c] - c1[2 * c];
                        dots[2 * c + 1] = 2.0 * dots[2 * c + 1] - c1[2 * c + 1];
                    }
                    col_mean(R, dots, mu + 2 * (2 * n_ - 1));
                    STORE_COLS(2 * n_ - 1, dots);
                }
                if (2 * n_ < M) {
                    v_dot(rows, R, t1, t1, dots);
                    for (int32_t c = 0; c < R; ++c) {
                        dots[2 * c] = 2.0 * dots[2 * c] - c0[2 * c];
                        dots[2 * c + 1] = 2.0 * dots[2 * c + 1] - c0[2 * c + 1];
                    }
                    col_mean(R, dots, mu + 2 * (2 * n_));
                    STORE_COLS(2 * n_, dots);
                }
            }
            if (2 * n_ + 1 >= M) break;
            if (!fused) {
                or_apply_op(op, t1, R, t2, tmp);
                v_axpby(n, -1.0, t0, 2.0, t2); /* t2 = 2 t2 - t0 (:170) */
            } else {
                /* av = A t1 (unscaled), then the fused epilogue + dots */
                if (op->form == OR_FORM_ORTHOGONAL) {
                    or_bsr_mult(op->H, t1, R, tmp);
                } else {
                    or_bsr_mult(op->Z, t1, R, t2);
                    or_bsr_mult(op->H, t2, R, tmp);
                    if (op->ZH) or_bsr_mult(op->ZH, tmp, R, t2);
                    else or_bsr_mult_adjoint(op->Z, tmp, R, t2);
                    memcpy(tmp, t2, sizeof(double) * 2 * (size_t)n);
                }
                fused_step(op, t0, t1, t2, tmp, R, dots, dots2);
                /* dots = <t2|t2> -> mu_{2n+2}; dots2 = <t2|t1> -> mu_{2n+1} */
                const int m_odd = 2 * (n_ + 1) - 1, m_even = 2 * (n_ + 1);
                if (m_odd < M) {
                    for (int32_t c = 0; c < R; ++c) {
                        dots2[2 * c] = 2.0 * dots2[2 * c] - c1[2 * c];
                        dots2[2 * c + 1] = 2.0 * dots2[2 * c + 1] - c1[2 * c + 1];
                    }
                    col_mean(R, dots2, mu + 2 * m_odd);
                    STORE_COLS(m_odd, dots2);
                }
                if (m_even < M) {
                    for (int32_t c = 0; c < R; ++c) {
                        dots[2 * c] = 2.0 * dots[2 * c] - c0[2 * c];
                        dots[2 * c + 1] = 2.0 * dots[2 * c + 1] - c0[2 * c + 1];
                    }
                    col_mean(R, dots, mu + 2 * m_even);
                    STORE_COLS(m_even, dots);
                }
            }
            double *sw = t0; t0 = t1; t1 = t2; t2 = sw; /* swaps (:171-172) */
        }
    } else if (M > 1) {
        for (int32_t m = 2; m < M; ++m) { /* plain recurrence (:175-187) */
            or_apply_op(op, t1, R, t2, tmp);
            v_axpby(n, -1.0, t0, 2.0, t2);
            v_dot(rows, R, psi, t2, dots);
            col_mean(R, dots, mu + 2 * m);
            STORE_COLS(m, dots);
            double *sw = t0; t0 = t1; t1 = t2; t2 = sw;
        }
    }
#undef STORE_COLS
    free(t0); free(t1); free(t2); free(tmp); free(dots); free(dots2); free(c0); free(c1);
    return 0;
}

/* ------------------------------------------------- projected moments --- */

/* pdos.hpp:185-202: raw[slot] += sum_c conj(a(row, c)) b(row, c), slot 0 the
 * all-rows channel, slot 1 + g for rows of group g >= 0. Serial over rows. */
static void seg_dot(int64_t rows, int32_t R, const double *a, const double *b,
                    const int32_t *grp, double *raw) {
    for (int64_t i = 0; i < rows; ++i) {
        double sr = 0.0, si = 0.0;
        const double *ar = a + 2 * i * R, *brw = b + 2 * i * R;
        for (int32_t c = 0; c < R; ++c) {
            sr += ar[2 * c] * brw[2 * c] + ar[2 * c + 1] * brw[2 * c + 1];
            si += ar[2 * c] * brw[2 * c + 1] - ar[2 * c + 1] * brw[2 * c];
        }
        raw[0] += sr;
        raw[1] += si;
        if (grp[i] >= 0) {
            raw[2 * (1 + (int64_t)grp[i])] += sr;
            raw[2 * (1 + (int64_t)grp[i]) + 1] += si;
        }
    }
}

/* kpm_pdos_moments multivector (pdos.hpp:149-266): the unfused recurrence of
 * or_kpm_moments with the segmented dots, one reduction at the end, the
 * division by R, then the per-channel doubling transform (:236-263).
 *   group_of_row: [rows], ids < G or negative (no group).
 *   total: [M] complex. group: [G][M] complex.
 * Returns 0; -1 on bad arguments; -2 when a group id is >= G (:163-167). */
int or_kpm_pdos_moments(const or_op *op, const double *psi, int32_t R, int32_t M,
                        int32_t doubling, int32_t G, const int32_t *group_of_row, double *total,
                        double *group) {
    if (M < 1 || R < 1 || G < 0) return -1;
    const int64_t rows = op->H->nbr * op->H->b;
    for (int64_t i = 0; i < rows; ++i)
        if (group_of_row[i] >= G) return -2;
    const int64_t n = rows * (int64_t)R;
    const int64_t width = 1 + (int64_t)G;
    double *raw = (double *)calloc(2 * (size_t)M * (size_t)width, sizeof(double));
    double *t0 = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    double *t1 = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    double *t2 = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    double *tmp = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    if (!raw || !t0 || !t1 || !t2 || !tmp) return -1;
#define RAW(m) (raw + 2 * (size_t)(m) * (size_t)width)
    memcpy(t0, psi, sizeof(double) * 2 * (size_t)n);
    seg_dot(rows, R, t0, t0, group_of_row, RAW(0));
    if (M > 1) {
        or_apply_op(op, t0, R, t1, tmp);
        seg_dot(rows, R, t0, t1, group_of_row, RAW(1));
        if (doubling) {
            for (int n_ = 1;; ++n_) {
                if (2 * n_ - 1 < M && 2 * n_ - 1 > 1) seg_dot(rows, R, t1, t0, group_of_row, RAW(2 * n_ - 1));
                if (2 * n_ < M) seg_dot(rows, R, t1, t1, group_of_row, RAW(2 * n_));
                if (2 * n_ + 1 >= M) break;
                or_apply_op(op, t1, R, t2, tmp);
                v_axpby(n, -1.0, t0, 2.0, t2);
                double *sw = t0; t0 = t1; t1 = t2; t2 = sw;
            }
        } else {
            for (int32_t m = 2; m < M; ++m) {
                or_apply_op(op, t1, R, t2, tmp);
                v_axpby(n, -1.0, t0, 2.0, t2);
                seg_dot(rows, R, psi, t2, group_of_row, RAW(m));
                double *sw = t0; t0 = t1; t1 = t2; t2 = sw;
            }
        }
    }
    const double inv = 1.0 / (double)R;
    for (size_t k = 0; k < 2 * (size_t)M * (size_t)width; ++k) raw[k] *= inv;
    for (int64_t ch = 0; ch < width; ++ch) {
        double *out = ch == 0 ? total : group + 2 * (size_t)(ch - 1) * (size_t)M;
        const double *r0 = RAW(0) + 2 * ch;
        const double *r1 = M > 1 ? RAW(1) + 2 * ch : NULL;
        for (int32_t m = 0; m < M; ++m) {
            const double *rm = RAW(m) + 2 * ch;
            double re = rm[0], im = rm[1];
            if (doubling && m >= 2) {
                const double *base = (m % 2 == 0) ? r0 : r1;
                re = 2.0 * re - base[0];
                im = 2.0 * im - base[1];
            }
            out[2 * m] = re;
            out[2 * m + 1] = im;
        }
    }
#undef RAW
    free(raw); free(t0); free(t1); free(t2); free(tmp);
    return 0;
}

/* ---------------------------------------------------------- kernels --- */

/* kpm.hpp:18-26 */
void or_jackson_kernel(int32_t n, double *out) {
    const double denom = (double)(n + 1);
    const double theta = M_PI / denom;
    const double cot_theta = cos(theta) / sin(theta);
    out[0] = 1.0;
    for (int32_t i = 1; i < n; ++i)
        out[i] = ((denom - i) * cos(theta * i) + cot_theta * sin(theta * i)) / denom;
}

/* kpm.hpp:52-85; kernel 0 none, 1 jackson, 2 cos taper, 3 vp taper. */
int or_kpm_kernel_weights(int32_t kernel, int32_t n, double *g) {
    if (n <= 0) return -1;
    for (int32_t i = 0; i < n; ++i) g[i] = 1.0;
    const int32_t n0 = n / 2;
    switch (kernel) {
    case 0: break;
    case 1: or_jackson_kernel(n, g); break;
    case 2:
        for (int32_t i = n0; i < n; ++i) {
            const double s = (double)(i - n0) / (double)(n - n0);
            g[i] = 0.5 * (1.0 + cos(M_PI * s));
        }
        break;
    case 3:
        for (int32_t i = n0; i < n; ++i) g[i] = (double)(n - i) / (double)(n - n0);
        break;
    default: return -1;
    }
    return 0;
}

/* kpm.hpp:206-241: rho(x) = (g0 mu0 + 2 sum g_n Re mu_n T_n(x)) / (pi sqrt(max(1e-14, 1-x^2)))
 * with x clamped to [-1 + inset, 1 - inset]. mu: [M] complex interleaved. */
int or_kpm_density(const double *mu, int32_t M, const double *x, int64_t nx, double edge_inset,
                   int32_t kernel, double *rho) {
    if (M == 0) {
        for (int64_t j = 0; j < nx; ++j) rho[j] = 0.0;
        return 0;
    }
    double *g = (double *)malloc(sizeof(double) * (size_t)M);
    if (!g || or_kpm_kernel_weights(kernel, M, g) != 0) { free(g); return -1; }
    for (int64_t j = 0; j < nx; ++j) {
        double xv = x[j];
        const double lo = -1.0 + edge_inset, hi = 1.0 - edge_inset;
        if (xv > hi) xv = hi;
        if (xv < lo) xv = lo;
        double t_prev = 1.0, t_curr = xv;
        double acc = g[0] * mu[0];
        if (M > 1) acc += 2.0 * g[1] * mu[2] * t_curr;
        for (int32_t m = 2; m < M; ++m) {
            const double t_next = 2.0 * xv * t_curr - t_prev;
            acc += 2.0 * g[m] * mu[2 * m] * t_next;
            t_prev = t_curr;
            t_curr = t_next;
        }
        const double w = 1.0 - xv * xv;
        rho[j] = acc / (M_PI * sqrt(w > 1e-14 ? w : 1e-14));
    }
    free(g);
    return 0;
}

int or_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void or_set_num_threads(int32_t n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}
