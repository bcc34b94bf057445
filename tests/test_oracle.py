"""Pins the CPU oracle (oracle/kpm_oracle.c) before it is trusted as the checker.

Every check here comes from outside the oracle's own formulas: the
reference's own known-answer tests (restated, paths relative to the RSATB
tree), published test vectors, closed forms, and dense LAPACK references on
our own matrices. Each is chosen so that a plausible slip in the oracle (a
dropped term, a wrong sign or index, a transposed operand) fails it.
"""
import math

import numpy as np
import pytest

from rsatb_amd.bsr import BSR
from rsatb_amd.lattice import (diamond_supercell, graphene_primitive_cell, graphene_rect_supercell,
                               neighbor_edges, tile)
from rsatb_amd.models import RandomBlockModel, SkPzModel, assemble


def diag_bsr(vals):
    n = len(vals)
    return BSR.from_blocks(n, 1, np.arange(n), np.arange(n), np.asarray(vals, dtype=np.complex128).reshape(-1, 1, 1))


def random_hermitian_bsr(n_blocks, b, rng, density=0.4, complex_=True):
    rows, cols, blks = [], [], []
    for i in range(n_blocks):
        for j in range(i, n_blocks):
            if i != j and rng.random() > density:
                continue
            B = rng.standard_normal((b, b)) + (1j * rng.standard_normal((b, b)) if complex_ else 0)
            if i == j:
                B = 0.5 * (B + B.conj().T)
                rows.append(i); cols.append(i); blks.append(B)
            else:
                rows += [i, j]; cols += [j, i]; blks += [B, B.conj().T]
    return BSR.from_blocks(n_blocks, b, np.array(rows), np.array(cols), np.array(blks))


# ------------------------------------------------------------------ RNG
def test_splitmix64_published_vector(oracle):
    # SplitMix64 seeded with 0 (Steele, Lea & Flood 2014): the first outputs
    # are mix(k * gamma) for k = 1.. with the finaliser adding gamma first.
    gamma = 0x9E3779B97F4A7C15
    expected = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F, 0xF88BB8A8724C81EC]
    got = [oracle.mix64((k * gamma) & (2**64 - 1)) for k in range(4)]
    assert got == expected


def test_rademacher_sign_statistics(oracle):
    # rs_ldos.hpp:249-264: +-1, key-pure, balanced, decorrelated over columns
    s = np.sign(oracle.samples_rademacher(20240917, 4000, 8).real)
    assert np.allclose(np.abs(oracle.samples_rademacher(20240917, 4000, 8)), 1.0 / math.sqrt(4000), rtol=0, atol=0)
    assert set(np.unique(s)) == {-1.0, 1.0}
    assert abs(s.mean()) < 4.0 / math.sqrt(s.size)
    c = np.corrcoef(s.T)
    assert np.max(np.abs(c - np.eye(8))) < 0.08
    assert oracle.rademacher_sign(20240917, 7, 3) == s[7, 3]
    # a different seed gives a different sign pattern
    s2 = oracle.samples_rademacher(20240918, 4000, 8).real
    assert np.mean(np.sign(s2) == np.sign(s)) < 0.6


def test_phase_samples_unit_norm(oracle):
    p = oracle.samples_phase(7, 1000, 5)
    assert np.allclose(np.abs(p) ** 2 * 1000, 1.0, atol=1e-12)
    assert np.allclose(np.sum(np.abs(p) ** 2, axis=0), 1.0, atol=1e-12)
    ang = np.angle(p[:, 0])
    # uniform phase: first circular moment ~ 0
    assert abs(np.mean(np.exp(1j * ang))) < 4.0 / math.sqrt(1000)


# ---------------------------------------------------------- BSR kernels
@pytest.mark.parametrize("b", [1, 2, 3])
def test_bsr_mult_and_adjoint_against_dense(oracle, b):
    rng = np.random.default_rng(b)
    nbr = 7
    rows, cols = np.nonzero(rng.random((nbr, nbr)) < 0.4)
    # non-symmetric blocks pin row-major storage (test_preconditioner.cpp:37-80)
    blks = rng.standard_normal((rows.size, b, b)) + 1j * rng.standard_normal((rows.size, b, b))
    A = BSR.from_blocks(nbr, b, rows, cols, blks)
    D = A.to_dense()
    X = rng.standard_normal((nbr * b, 5)) + 1j * rng.standard_normal((nbr * b, 5))
    assert np.allclose(oracle.bsr_mult(A, X), D @ X, atol=1e-12, rtol=0)
    assert np.allclose(oracle.bsr_mult_adjoint(A, X), D.conj().T @ X, atol=1e-12, rtol=0)


def test_row_major_within_block(oracle):
    # one 2x2 block [[1, 2], [3, 4]] at (0, 0): y = A e_1 must be (2, 4)
    A = BSR.from_blocks(1, 2, np.array([0]), np.array([0]), np.array([[[1.0, 2.0], [3.0, 4.0]]]))
    y = oracle.bsr_mult(A, np.array([0.0, 1.0]))
    assert np.allclose(y[:, 0], [2.0, 4.0])


# ------------------------------------------------------------ apply_form
def test_sandwich_apply_exact(oracle):
    # test_cheby.cpp:400-470: H = diag(1..4), Z = 1/4 (block Jacobi of S = 4),
    # window [1/16, 4/16]: y_i = (h_i/16 - center)/half_width for x = 1
    H = diag_bsr([1.0, 2.0, 3.0, 4.0])
    Z = diag_bsr([0.25] * 4)
    emin, emax = 1.0 / 16.0, 4.0 / 16.0
    c, a = (emax + emin) / 2.0, (emax - emin) / 2.0
    op = oracle.Operator(oracle.FORM_SANDWICH, H, c, a, Z)
    y = op.apply(np.ones(4))[:, 0]
    expected = (np.array([1.0, 2.0, 3.0, 4.0]) / 16.0 - c) / a
    assert np.max(np.abs(y - expected)) < 1e-12


def test_adjoint_identity_forms(oracle):
    # test_cheby.cpp:148-241: complex Hermitian tridiagonal H, a NON-symmetric
    # Z, so a sandwich that applied Z where Z^H belongs breaks Hermiticity.
    n = 6
    rows, cols, vals = [], [], []
    for i in range(n):
        rows.append(i); cols.append(i); vals.append(2.0 + 0.3 * i)
        if i > 0:
            rows.append(i); cols.append(i - 1); vals.append(0.4 + 0.7j)
        if i + 1 < n:
            rows.append(i); cols.append(i + 1); vals.append(0.4 - 0.7j)
    H = BSR.from_blocks(n, 1, np.array(rows), np.array(cols), np.array(vals, dtype=np.complex128).reshape(-1, 1, 1))
    rng = np.random.default_rng(3)
    zr, zc = np.nonzero(np.abs(np.subtract.outer(np.arange(n), np.arange(n))) <= 1)
    Z = BSR.from_blocks(n, 1, zr, zc, (rng.standard_normal(zr.size) + 1j * rng.standard_normal(zr.size)).reshape(-1, 1, 1))
    x = np.array([0.3 + i - (0.7 - 0.2 * i) * 1j for i in range(n)])
    y = np.array([-1.1 + 0.5 * i + (0.4 - 0.3 * i) * 1j for i in range(n)])
    for form, zz in [(oracle.FORM_ORTHOGONAL, None), (oracle.FORM_SANDWICH, Z)]:
        op = oracle.Operator(form, H, 0.1, 4.0, zz)
        lhs = np.vdot(y, op.apply(x)[:, 0])
        rhs = np.vdot(op.apply(y)[:, 0], x)
        assert abs(lhs - rhs) / abs(lhs) < 1e-12
    # the dense operator itself
    A = oracle.dense_scaled(oracle.FORM_SANDWICH, H, 0.1, 4.0, Z)
    assert np.allclose(oracle.Operator(oracle.FORM_SANDWICH, H, 0.1, 4.0, Z).apply(x)[:, 0], A @ x, atol=1e-13)


# ------------------------------------------------------------- moments
def two_level_fixture(n=100):
    # test_dos.cpp:69-96: H = diag(+-0.5), samples ((7g+3)%5<2 ? 1 : -1)/sqrt(N)
    H = diag_bsr([0.5 if g % 2 == 0 else -0.5 for g in range(n)])
    psi = np.array([(1.0 if (g * 7 + 3) % 5 < 2 else -1.0) / math.sqrt(n) for g in range(n)])
    return H, psi


def test_dos_fixture_analytic_moments(oracle):
    # test_dos.cpp:101-139: mu_n = cos(n pi/3) (even n), 0 (odd), 1e-10; doubled == plain
    H, psi = two_level_fixture()
    op = oracle.Operator(oracle.FORM_ORTHOGONAL, H, 0.0, 1.0)
    M = 64
    mu = op.moments(psi, M, doubling=True)
    mu_plain = op.moments(psi, M, doubling=False)
    exact = np.array([math.cos(n * math.pi / 3.0) if n % 2 == 0 else 0.0 for n in range(M)])
    assert np.max(np.abs(mu - exact)) < 1e-10
    assert np.max(np.abs(mu - mu_plain)) < 1e-10
    # Jackson density: integral mu_0 = 1 within 0.02, positive peak at 0.5 +- 0.05
    ng = 400
    x = -1.0 + 2.0 * (np.arange(ng) + 0.5) / ng
    rho = oracle.kpm_density(mu, x, 1.0 / ng)
    assert abs(rho.sum() * 2.0 / ng - 1.0) < 0.02
    xp = x[x > 0][np.argmax(rho[x > 0])]
    assert abs(xp - 0.5) < 0.05


def test_dos_fixture_multivector(oracle):
    # test_dos.cpp:145-159: 4 random +-1/sqrt(N) columns; on this H the
    # estimator is exact per column, so the column mean is exact too
    H, _ = two_level_fixture()
    rng = np.random.default_rng(12345)
    psi = np.where(rng.integers(0, 2, (100, 4)) == 0, 1.0, -1.0) / 10.0
    op = oracle.Operator(oracle.FORM_ORTHOGONAL, H, 0.0, 1.0)
    mu = op.moments(psi, 40)
    exact = np.array([math.cos(n * math.pi / 3.0) if n % 2 == 0 else 0.0 for n in range(40)])
    assert np.max(np.abs(mu - exact)) < 1e-10


@pytest.mark.parametrize("form", [0, 1])
@pytest.mark.parametrize("fused", [False, True])
def test_moments_against_dense_eigendecomposition(oracle, form, fused):
    rng = np.random.default_rng(11 + form)
    b, nbr = 3, 8
    H = random_hermitian_bsr(nbr, b, rng)
    Z = None
    if form == 1:
        zr, zc = np.nonzero(rng.random((nbr, nbr)) < 0.5)
        Z = BSR.from_blocks(nbr, b, np.r_[zr, np.arange(nbr)], np.r_[zc, np.arange(nbr)],
                            np.concatenate([0.2 * rng.standard_normal((zr.size, b, b)),
                                            np.broadcast_to(np.eye(b), (nbr, b, b))]))
    A0 = oracle.dense_scaled(form, H, 0.0, 1.0, Z)
    ev = np.linalg.eigvalsh(0.5 * (A0 + A0.conj().T))
    emin, emax = ev[0] - 0.3, ev[-1] + 0.5
    c, a = (emax + emin) / 2.0, (emax - emin) / 2.0
    op = oracle.Operator(form, H, c, a, Z)
    psi = oracle.samples_phase(5, nbr * b, 6)
    M = 50
    ref = oracle.dense_moments(oracle.dense_scaled(form, H, c, a, Z), psi, M)
    for dbl in (True, False):
        mu = op.moments(psi, M, doubling=dbl, fused=fused)
        assert np.max(np.abs(mu - ref)) < 1e-11, (dbl, fused)


def test_moment_invariants_plant_detects_sign_slip(oracle):
    # mu(-H) = (-1)^n mu(H): a sign error in the recurrence breaks it
    rng = np.random.default_rng(4)
    H = random_hermitian_bsr(10, 2, rng)
    Hn = BSR(H.n_block_rows, H.block_dim, H.row_ptr, H.col_ind, -H.values)
    psi = oracle.samples_phase(9, 20, 3)
    a = 1.2 * np.max(np.abs(np.linalg.eigvalsh(H.to_dense())))
    m1 = oracle.Operator(0, H, 0.0, a).moments(psi, 30)
    m2 = oracle.Operator(0, Hn, 0.0, a).moments(psi, 30)
    sgn = (-1.0) ** np.arange(30)
    assert np.max(np.abs(m2 - sgn * m1)) < 1e-12
    assert abs(m1[0] - 1.0) < 1e-14
    assert np.all(np.abs(m1) <= 1.0 + 1e-12)


# ----------------------------------------------------------------- KPM
def test_jackson_closed_forms(oracle):
    g = oracle.jackson_kernel(10)
    assert g[0] == 1.0
    N = 10
    n = np.arange(N)
    ref = ((N + 1 - n) * np.cos(np.pi * n / (N + 1)) + np.sin(np.pi * n / (N + 1)) / np.tan(np.pi / (N + 1))) / (N + 1)
    ref[0] = 1.0
    assert np.allclose(g, ref, atol=1e-14)
    # Jackson weights are positive and decreasing for this definition
    assert np.all(np.diff(g) < 0) and np.all(g > 0)
    # tapers: first half untouched (kpm.hpp:63-82)
    for k in (2, 3):
        w = oracle.kernel_weights(k, 8)
        assert np.all(w[:4] == 1.0) and w[4] == 1.0 and w[-1] < 1.0


def test_density_integrates_to_mu0(oracle):
    # int rho dx = mu_0 (kpm.hpp:198-201): exact Gauss-Chebyshev quadrature
    rng = np.random.default_rng(2)
    M = 40
    mu = rng.standard_normal(M) * 0.3
    mu[0] = 0.8
    K = 200
    x = np.cos(np.pi * (np.arange(K) + 0.5) / K)
    rho = oracle.kpm_density(mu, x, 1e-12)
    integral = np.pi / K * np.sum(rho * np.sqrt(1 - x * x))
    assert abs(integral - 0.8) < 1e-12


def test_jackson_fwhm_matches_resolution(oracle):
    # observables.hpp:337-350: a delta spectrum broadens to FWHM 7.396 a / N
    N = 400
    x0 = 0.1
    mu = np.cos(np.arange(N) * np.arccos(x0))
    x = np.linspace(-0.2, 0.4, 20001)
    rho = oracle.kpm_density(mu, x, 1e-9)
    half = rho.max() / 2.0
    above = x[rho >= half]
    fwhm = above[-1] - above[0]
    # in scaled units (a = 1) the FWHM is sqrt(8 ln 2) pi / N = 7.396 / N
    assert abs(fwhm * N - 7.396) / 7.396 < 0.01


def test_dos_plan_formula(oracle):
    # observables.hpp:336-350 with the graphene window +-8.6 and dE = 0.15
    nt, n = oracle.dos_plan(-8.6, 8.6, 0.15)
    assert nt == math.ceil(17.2 / 0.15)
    assert n == math.ceil(math.sqrt(8 * math.log(2)) * math.pi * 8.6 / 0.15)


# --------------------------------------------------- graph / structure
@pytest.mark.parametrize("reps", [1, 2, 3, 4, 5])
def test_graphene_supercell_bound_3t(reps):
    # test/test_supercell_bonds.py:229-259: |E|max = 3|t| = 8.1 exactly
    st = tile(graphene_primitive_cell(), (reps, reps, 1))
    H = assemble(st, neighbor_edges(st, 1.6), SkPzModel())
    ev = np.linalg.eigvalsh(H.to_dense())
    assert abs(np.max(np.abs(ev)) - 8.1) < 1e-6


def test_neighbor_counts():
    e = neighbor_edges(graphene_rect_supercell(6, 5), 1.6)
    assert np.all(np.bincount(e.src) == 3)
    e = neighbor_edges(diamond_supercell(3), 2.7)
    counts = np.bincount(e.src)
    assert np.all(counts == 16)
    assert np.sum(e.dist < 1.6) == 4 * 216


def test_isqrt_patches_exact_cases(oracle):
    # (1) block-diagonal S: Z = S_ii^{-1/2} block by block
    rng = np.random.default_rng(8)
    b = 3
    blks = []
    for _ in range(4):
        Q = rng.standard_normal((b, b))
        blks.append(Q @ Q.T + b * np.eye(b))
    S = BSR.from_blocks(4, b, np.arange(4), np.arange(4), np.array(blks))
    Z = oracle.isqrt_patches(S)
    for i, B in enumerate(blks):
        w, V = np.linalg.eigh(B)
        assert np.allclose(Z.values[i], (V / np.sqrt(w)) @ V.T, atol=1e-12)
    # (2) every row couples to every row: the patch is the whole S, Z = S^{-1/2}
    n = 4
    Q = rng.standard_normal((n * b, n * b)) * 0.1
    D = Q @ Q.T + np.eye(n * b)
    rows, cols = np.nonzero(np.ones((n, n)))
    Sb = BSR.from_blocks(n, b, rows, cols, np.array([D[i * b:(i + 1) * b, j * b:(j + 1) * b] for i, j in zip(rows, cols)]))
    Zd = oracle.isqrt_patches(Sb).to_dense()
    from scipy.linalg import fractional_matrix_power
    assert np.allclose(Zd, np.real(fractional_matrix_power(D, -0.5)), atol=1e-10)
    assert np.allclose(Zd @ D @ Zd, np.eye(n * b), atol=1e-10)


def test_random_block_model_symmetric_spd():
    st = diamond_supercell(2)
    e = neighbor_edges(st, 2.7)
    m = RandomBlockModel()
    H = assemble(st, e, m, "H").to_dense()
    S = assemble(st, e, m, "S").to_dense()
    assert np.array_equal(H, H.T) and np.array_equal(S, S.T)
    assert np.linalg.eigvalsh(S)[0] > 0.3
    # blocks per row: self + 4 + 12
    Hb = assemble(st, e, m, "H")
    assert np.all(np.diff(Hb.row_ptr) == 17)


def test_dense_moment_gradient_against_finite_differences(oracle):
    # the Daleckii-Krein reference of the adjoint pass, pinned by central
    # differences of eigendecomposition moments (a route with no recurrence)
    rng = np.random.default_rng(0)
    n = 12
    B = rng.standard_normal((n, n))
    A = (B + B.T) / 2
    A /= 1.2 * np.max(np.abs(np.linalg.eigvalsh(A)))
    psi = rng.standard_normal((n, 3)) + 1j * rng.standard_normal((n, 3))
    psi /= np.linalg.norm(psi, axis=0)
    M = 20
    g = rng.standard_normal(M)
    G = oracle.dense_moment_gradient(A, psi, g)
    assert np.allclose(G, G.T, atol=1e-12)

    def loss(X):
        return float(np.sum(g * oracle.dense_moments(X, psi, M).real))

    h = 1e-6
    for i, j in [(0, 0), (1, 2), (5, 7), (11, 3)]:
        E = np.zeros((n, n))
        E[i, j] = E[j, i] = 1.0 if i != j else 1.0
        fd = (loss(A + h * E) - loss(A - h * E)) / (2 * h)
        assert abs(fd - (G[i, j] + G[j, i] if i != j else G[i, i])) < 1e-7


def test_recurrence_gradient_matches_dense(oracle):
    # central differences of the oracle's doubled moments under a symmetric
    # perturbation of one stored block pair agree with the dense reference
    # (the per-entry derivative of the doubled recurrence is defined only up
    # to its symmetric part: G_ij + G_ji is what a symmetric H sees)
    rng = np.random.default_rng(3)
    H = random_hermitian_bsr(10, 2, rng, complex_=False)
    H = BSR(H.n_block_rows, 2, H.row_ptr, H.col_ind, H.values.real.copy())
    a = 1.1 * np.max(np.abs(np.linalg.eigvalsh(H.to_dense())))
    psi = oracle.samples_phase(3, 20, 4)
    M = 24
    g = rng.standard_normal(M)
    G = oracle.dense_moment_gradient(H.to_dense() / a, psi, g)
    rows = H.block_rows()
    s = int(np.flatnonzero(rows != H.col_ind)[0])  # an off-diagonal block
    r, cidx = int(rows[s]), int(H.col_ind[s])
    s2 = int(np.flatnonzero((rows == cidx) & (H.col_ind == r))[0])
    h = 1e-6
    for (p, q) in [(0, 0), (0, 1), (1, 0)]:
        lp_lm = []
        for sign in (1.0, -1.0):
            Hx = BSR(H.n_block_rows, 2, H.row_ptr, H.col_ind, H.values.copy())
            Hx.values[s, p, q] += sign * h
            Hx.values[s2, q, p] += sign * h
            lp_lm.append(float(np.sum(g * oracle.Operator(0, Hx, 0.0, a).moments(psi, M).real)))
        fd = (lp_lm[0] - lp_lm[1]) / (2 * h)
        ref = (G[2 * r + p, 2 * cidx + q] + G[2 * cidx + q, 2 * r + p]) / a
        assert abs(fd - ref) < 1e-7


# ---------------------------------------------- gradient through S^{-1/2}
def _spd_overlap(n_blocks, b, rng, scale=0.08):
    """S = I + scale * (random symmetric block pattern), SPD by construction
    (diagonally dominant for this scale)."""
    Hr = random_hermitian_bsr(n_blocks, b, rng, complex_=False)
    vals = scale * Hr.values.real / np.max(np.abs(Hr.values.real))
    rows = Hr.block_rows()
    for s in np.flatnonzero(rows == Hr.col_ind):
        vals[s] = 0.5 * (vals[s] + vals[s].T) + np.eye(b)
    return BSR(n_blocks, b, Hr.row_ptr, Hr.col_ind, vals)


def _sym_direction(S, rng):
    """A random direction on S's pattern with D(j, i) = D(i, j)^T."""
    rows = S.block_rows()
    D = np.zeros_like(S.values)
    for s in range(S.values.shape[0]):
        r, c = int(rows[s]), int(S.col_ind[s])
        if r > c:
            continue
        B = rng.standard_normal(S.values.shape[1:])
        if r == c:
            D[s] = 0.5 * (B + B.T)
        else:
            s2 = int(np.flatnonzero((rows == c) & (S.col_ind == r))[0])
            D[s], D[s2] = B, B.T
    return D


@pytest.mark.parametrize("b", [1, 2])
def test_isqrt_patches_vjp_against_finite_differences(oracle, b):
    # the Daleckii-Krein patch VJP against central differences of the
    # eigen-route factor along symmetric pattern-preserving directions
    rng = np.random.default_rng(11 + b)
    S = _spd_overlap(9, b, rng)
    W = rng.standard_normal(S.values.shape)  # L = <W, Z(S)>
    dS = oracle.isqrt_patches_vjp(S, W)

    def loss(vals):
        return float(np.sum(W * oracle.isqrt_patches(BSR(S.n_block_rows, b, S.row_ptr, S.col_ind, vals)).values))

    h = 1e-5
    for _ in range(3):
        D = _sym_direction(S, rng)
        fd = (loss(S.values + h * D) - loss(S.values - h * D)) / (2 * h)
        assert abs(fd - float(np.sum(dS * D))) < 1e-8 * max(1.0, abs(fd))
    # a planted slip (the transposed derivative operand) is caught
    bad = oracle.isqrt_patches_vjp(S, np.swapaxes(W, 1, 2)) if b > 1 else -dS
    D = _sym_direction(S, rng)
    fd = (loss(S.values + h * D) - loss(S.values - h * D)) / (2 * h)
    assert abs(fd - float(np.sum(bad * D))) > 1e-6


def test_sandwich_gradients_against_finite_differences(oracle):
    # graphene p_z with onsite, three hopping shells and a nearest-neighbour
    # overlap: dL/dtheta from the dense chain (H and S gradients contracted
    # with the affine models) against central differences of the
    # eigendecomposition moments of Z(S)^T H Z(S)
    from rsatb_amd.autograd import sk_pz_overlap_parameterization, sk_pz_parameterization

    st = graphene_rect_supercell(3, 3)
    e = neighbor_edges(st, 2.9)
    hm = sk_pz_parameterization(st, e, [1.6, 2.5, 2.9])
    sm = sk_pz_overlap_parameterization(st, neighbor_edges(st, 1.6), [1.6])
    th_h = np.array([0.1, -2.7, -0.2, -0.18])
    th_s = np.array([0.1])
    psi = oracle.samples_rademacher(7, st.n_atoms, 3)
    M = 16
    g = np.random.default_rng(2).standard_normal(M)
    center, a = 0.3, 12.0
    out = oracle.sandwich_moment_gradients(hm.build(th_h), sm.build(th_s), center, a, psi, g)
    grad = np.concatenate([hm.contract(out["H"]), sm.contract(out["S"])])

    def loss(th):
        H, S = hm.build(th[:4]), sm.build(th[4:])
        Z = oracle.isqrt_patches(S).to_dense()
        A = (Z.T @ H.to_dense() @ Z - center * np.eye(st.n_atoms)) / a
        return float(np.sum(g * oracle.dense_moments(A, psi, M).real))

    th = np.concatenate([th_h, th_s])
    h = 1e-5
    for k in range(th.size):
        d = np.zeros(th.size)
        d[k] = h
        fd = (loss(th + d) - loss(th - d)) / (2 * h)
        assert abs(fd - grad[k]) < 1e-7 * max(1.0, abs(fd)), (k, fd, grad[k])
    assert abs(grad[4]) > 1e-3  # the overlap derivative is not trivially zero
