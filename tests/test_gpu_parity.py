"""GPU parity: the gfx950 library against the CPU oracle on the same seeded
inputs, through the C ABI. Integer/index work (samples' signs, structure) is
bit-exact; FP64 arithmetic agrees to 1e-12 relative per element and per
moment (the FP64 tolerance; summation order is the only difference)."""
import math
import threading

import numpy as np
import pytest

from rsatb_amd import _lib
from rsatb_amd.bsr import BSR
from rsatb_amd.lattice import diamond_supercell, graphene_rect_supercell, neighbor_edges
from rsatb_amd.models import RandomBlockModel, SkPzModel, assemble

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12


def rel_err(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def random_bsr(nbr, b, rng, density=0.3, complex_=False, square_diag=True):
    mask = rng.random((nbr, nbr)) < density
    if square_diag:
        np.fill_diagonal(mask, True)
    rows, cols = np.nonzero(mask)
    blks = rng.standard_normal((rows.size, b, b))
    if complex_:
        blks = blks + 1j * rng.standard_normal((rows.size, b, b))
    return BSR.from_blocks(nbr, b, rows, cols, blks)


def hermitian_bsr(nbr, b, rng, complex_=False, density=0.3):
    A = random_bsr(nbr, b, rng, density, complex_)
    D = A.to_dense()
    D = 0.5 * (D + D.conj().T)
    rows, cols = np.nonzero(np.abs(D).reshape(nbr, b, nbr, b).sum(axis=(1, 3)) > 0)
    blks = np.array([D[i * b:(i + 1) * b, j * b:(j + 1) * b] for i, j in zip(rows, cols)])
    return BSR.from_blocks(nbr, b, rows, cols, blks)


# ------------------------------------------------------------- samples
def test_rademacher_samples_bit_exact(gpu_ctx, oracle):
    for rows, R in [(1000, 16), (777, 5), (64, 64)]:
        gpu_ctx.set_samples(rows, R, _lib.SAMPLES_RADEMACHER, 20240917)
        g = gpu_ctx.get_samples(as_complex=False)
        o = oracle.samples_rademacher(20240917, rows, R)
        assert np.array_equal(g, o.real)


def test_phase_samples(gpu_ctx, oracle):
    gpu_ctx.set_samples(999, 7, _lib.SAMPLES_PHASE, 123)
    g = gpu_ctx.get_samples(as_complex=True)
    o = oracle.samples_phase(123, 999, 7)
    # the hash (integers) is exact; sin/cos differ at most by a few ulp
    assert np.max(np.abs(g - o)) <= 4 * np.finfo(float).eps / math.sqrt(999)


# ---------------------------------------------------------------- SpMM
@pytest.mark.parametrize("b", [1, 2, 3, 13])
@pytest.mark.parametrize("R", [1, 5, 16, 33, 64, 100, 128, 200])
@pytest.mark.parametrize("kind", ["real", "realmat_cvec", "complex"])
def test_bsr_mult_matches_oracle(gpu_ctx, oracle, b, R, kind):
    rng = np.random.default_rng(b * 1000 + R)
    nbr = 37 if b < 13 else 9
    A = random_bsr(nbr, b, rng, complex_=(kind == "complex"))
    X = rng.standard_normal((nbr * b, R))
    if kind != "real":
        X = X + 1j * rng.standard_normal((nbr * b, R))
    gpu_ctx.upload_bsr(_lib.SLOT_Z, A)
    y = gpu_ctx.bsr_mult(_lib.SLOT_Z, X)
    ref = oracle.bsr_mult(A, X)
    if kind == "real":
        assert not np.iscomplexobj(y)
    assert rel_err(y, ref) < FP64_TOL


def test_empty_rows_and_zero_blocks(gpu_ctx, oracle):
    # rows without stored blocks give zero rows; explicit zero blocks are fine
    nbr, b = 6, 2
    rows = np.array([0, 0, 2, 5, 5])
    cols = np.array([0, 3, 2, 1, 5])
    blks = np.random.default_rng(0).standard_normal((5, b, b))
    blks[2] = 0.0
    A = BSR.from_blocks(nbr, b, rows, cols, blks)
    X = np.random.default_rng(1).standard_normal((nbr * b, 9))
    gpu_ctx.upload_bsr(_lib.SLOT_H, A)
    y = gpu_ctx.bsr_mult(_lib.SLOT_H, X)
    assert np.array_equal(y[2:4], np.zeros((2, 9))) and np.array_equal(y[6:10], np.zeros((4, 9)))
    assert rel_err(y, oracle.bsr_mult(A, X).real) < FP64_TOL


# ------------------------------------------------------------ operator
@pytest.mark.parametrize("form", [_lib.FORM_ORTHOGONAL, _lib.FORM_SANDWICH])
@pytest.mark.parametrize("complex_h", [False, True])
def test_apply_op_matches_oracle(gpu_ctx, oracle, form, complex_h):
    rng = np.random.default_rng(7 + form + 2 * complex_h)
    nbr, b = 20, 3
    H = hermitian_bsr(nbr, b, rng, complex_h)
    Z = random_bsr(nbr, b, rng, 0.25) if form == _lib.FORM_SANDWICH else None
    gpu_ctx.upload_bsr(_lib.SLOT_H, H)
    if Z is not None:
        gpu_ctx.upload_bsr(_lib.SLOT_Z, Z)
        gpu_ctx.upload_bsr(_lib.SLOT_ZH, Z.conj_transpose())
    gpu_ctx.set_operator(form, 0.3, 2.5)
    x = rng.standard_normal((nbr * b, 6)) + 1j * rng.standard_normal((nbr * b, 6))
    y = gpu_ctx.apply_op(x)
    ref = oracle.Operator(form, H, 0.3, 2.5, Z).apply(x)
    assert rel_err(y, ref) < FP64_TOL


def test_sandwich_apply_exact_fixture(gpu_ctx):
    # test_cheby.cpp:400-470 on the device
    H = BSR.from_blocks(4, 1, np.arange(4), np.arange(4), np.arange(1.0, 5.0).reshape(-1, 1, 1))
    Z = BSR.from_blocks(4, 1, np.arange(4), np.arange(4), np.full((4, 1, 1), 0.25))
    emin, emax = 1.0 / 16.0, 4.0 / 16.0
    c, a = (emax + emin) / 2.0, (emax - emin) / 2.0
    gpu_ctx.upload_bsr(_lib.SLOT_H, H)
    gpu_ctx.upload_bsr(_lib.SLOT_Z, Z)
    gpu_ctx.upload_bsr(_lib.SLOT_ZH, Z.conj_transpose())
    gpu_ctx.set_operator(_lib.FORM_SANDWICH, c, a)
    y = gpu_ctx.apply_op(np.ones(4))[:, 0]
    assert np.max(np.abs(y - (np.arange(1.0, 5.0) / 16.0 - c) / a)) < 1e-12


# ------------------------------------------------------------- moments
def test_dos_fixture_on_device(gpu_ctx):
    # test_dos.cpp:69-139: mu_n = cos(n pi/3) for even n, 0 for odd, 1e-10
    n = 100
    H = BSR.from_blocks(n, 1, np.arange(n), np.arange(n),
                        np.array([0.5 if g % 2 == 0 else -0.5 for g in range(n)]).reshape(-1, 1, 1))
    psi = np.array([(1.0 if (g * 7 + 3) % 5 < 2 else -1.0) / math.sqrt(n) for g in range(n)])
    gpu_ctx.upload_bsr(_lib.SLOT_H, H)
    gpu_ctx.set_operator(_lib.FORM_ORTHOGONAL, 0.0, 1.0)
    gpu_ctx.set_samples(n, 1, _lib.SAMPLES_UPLOAD_REAL, data=psi)
    exact = np.array([math.cos(k * math.pi / 3.0) if k % 2 == 0 else 0.0 for k in range(64)])
    for dbl in (_lib.DOUBLING_ON, _lib.DOUBLING_OFF):
        mu = gpu_ctx.kpm_moments(64, doubling=dbl)
        assert np.max(np.abs(mu - exact)) < 1e-10


@pytest.mark.parametrize("M", [1, 2, 3, 4, 5, 64, 65])
def test_moment_counts_edge_cases(gpu_ctx, oracle, M):
    rng = np.random.default_rng(M)
    H = hermitian_bsr(12, 2, rng)
    gpu_ctx.upload_bsr(_lib.SLOT_H, H)
    a = 1.1 * np.max(np.abs(np.linalg.eigvalsh(H.to_dense())))
    gpu_ctx.set_operator(_lib.FORM_ORTHOGONAL, 0.0, a)
    gpu_ctx.set_samples(24, 3, _lib.SAMPLES_PHASE, 5)
    psi = gpu_ctx.get_samples()
    op = oracle.Operator(0, H, 0.0, a)
    for dbl in (True, False):
        mu = gpu_ctx.kpm_moments(M, doubling=_lib.DOUBLING_ON if dbl else _lib.DOUBLING_OFF)
        ref = op.moments(psi, M, doubling=dbl)
        assert mu.shape == (M,)
        assert np.max(np.abs(mu - ref)) < 1e-12


def test_graphene_moments_and_dos_match_oracle(gpu_ctx, oracle):
    # graphene 16x16 rectangular (1024 atoms), R = 16 Rademacher, window +-8.1/0.99
    st = graphene_rect_supercell(16, 16)
    H = assemble(st, neighbor_edges(st, 1.6), SkPzModel())
    a = 8.1 / 0.99
    gpu_ctx.upload_bsr(_lib.SLOT_H, H)
    gpu_ctx.set_operator(_lib.FORM_ORTHOGONAL, 0.0, a)
    gpu_ctx.set_samples(st.n_atoms, 16, _lib.SAMPLES_RADEMACHER, 20240917)
    psi = oracle.samples_rademacher(20240917, st.n_atoms, 16)
    M = 256
    mu, cols = gpu_ctx.kpm_moments(M, per_column=True)
    ref, ref_cols = oracle.Operator(0, H, 0.0, a).moments(psi, M, per_column=True)
    assert np.max(np.abs(mu - ref)) < 1e-12
    assert np.max(np.abs(cols - ref_cols)) < 1e-12
    assert abs(mu[0] - 1.0) < 1e-14
    dos, omega = gpu_ctx.compute_dos(-a, a, 0.1)
    rdos, romega = oracle.compute_dos(oracle.Operator(0, H, 0.0, a), psi, -a, a, 0.1)
    assert np.array_equal(omega, romega)
    assert rel_err(dos, rdos) < 1e-10


@pytest.mark.parametrize("R", [16, 64, 128])
def test_diamond_sandwich_moments_match_oracle(gpu_ctx, oracle, R):
    # diamond 2x2x2 conventional (64 atoms, 13 orbitals), Z from the oracle's
    # patches, random phase, doubled and plain; R = 16, 64, 128 take the VALU
    # kernel and the MFMA kernels with 8 and 16 column tiles per wave. The
    # precision form follows the same moments to the FP32 bound.
    st = diamond_supercell(2)
    e = neighbor_edges(st, 2.7)
    m = RandomBlockModel()
    H = assemble(st, e, m, "H")
    S = assemble(st, e, m, "S")
    Z = oracle.isqrt_patches(S)
    Zb = BSR(Z.n_block_rows, Z.block_dim, Z.row_ptr, Z.col_ind, Z.values)
    A = oracle.dense_scaled(1, H, 0.0, 1.0, Z)
    ev = np.linalg.eigvalsh(0.5 * (A + A.T))
    emin, emax = ev[0] - 0.5, ev[-1] + 0.5
    c, a = (emax + emin) / 2, (emax - emin) / 2
    gpu_ctx.upload_bsr(_lib.SLOT_H, H)
    gpu_ctx.upload_bsr(_lib.SLOT_Z, Zb)
    gpu_ctx.upload_bsr(_lib.SLOT_ZH, Zb.conj_transpose())
    gpu_ctx.set_operator(_lib.FORM_SANDWICH, c, a)
    rows = st.n_atoms * 13
    gpu_ctx.set_samples(rows, R, _lib.SAMPLES_PHASE, 20240917)
    psi = gpu_ctx.get_samples()
    op = oracle.Operator(1, H, c, a, Z)
    M = 96
    for dbl in (True, False):
        d = _lib.DOUBLING_ON if dbl else _lib.DOUBLING_OFF
        mu = gpu_ctx.kpm_moments(M, doubling=d)
        ref = op.moments(psi, M, doubling=dbl)
        assert np.max(np.abs(mu - ref)) < 1e-12, dbl
        mu32 = gpu_ctx.kpm_moments(M, _lib.PREC_FP32, d)
        assert np.max(np.abs(mu32 - ref)) < 2e-5, dbl
    assert np.all(np.abs(mu) <= 1.0 + 1e-12)


def test_rows_mismatch_and_missing_factors(gpu_ctx):
    H = BSR.from_blocks(4, 1, np.arange(4), np.arange(4), np.ones((4, 1, 1)))
    gpu_ctx.upload_bsr(_lib.SLOT_H, H)
    gpu_ctx.set_operator(_lib.FORM_ORTHOGONAL, 0.0, 2.0)
    gpu_ctx.set_samples(5, 2, _lib.SAMPLES_RADEMACHER, 1)
    with pytest.raises(_lib.RsatbGpuError, match="differ"):
        gpu_ctx.kpm_moments(8)
    with pytest.raises(_lib.RsatbGpuError):
        gpu_ctx.set_samples(4, 300, _lib.SAMPLES_RADEMACHER, 1)
    with pytest.raises(_lib.RsatbGpuError):
        gpu_ctx.set_operator(_lib.FORM_ORTHOGONAL, 0.0, 0.0)


def test_cancel_between_launches(oracle):
    # the host flag is checked between steps; every launched kernel finishes
    ctx = _lib.Context(0)
    try:
        st = graphene_rect_supercell(32, 32)
        H = assemble(st, neighbor_edges(st, 1.6), SkPzModel())
        ctx.upload_bsr(_lib.SLOT_H, H)
        ctx.set_operator(_lib.FORM_ORTHOGONAL, 0.0, 8.2)
        ctx.set_samples(st.n_atoms, 16, _lib.SAMPLES_RADEMACHER, 3)
        t = threading.Timer(0.05, ctx.cancel)
        t.start()
        with pytest.raises(KeyboardInterrupt):
            ctx.kpm_moments(200_000)
        t.join()
        # the context is usable afterwards
        mu = ctx.kpm_moments(8)
        assert abs(mu[0] - 1.0) < 1e-14
    finally:
        ctx.close()


@pytest.mark.parametrize("kind", ["real", "realmat_cvec"])
@pytest.mark.parametrize("R", [1, 16, 33, 128])
def test_b1_empty_and_long_rows(gpu_ctx, oracle, kind, R):
    # scalar entries (the b = 1 kernel, real matrix, real or complex vectors):
    # empty rows, rows of every length up to dense rows of 300 entries (past
    # the 4-entry gather groups), every lane layout of R
    rng = np.random.default_rng(R + (7 if kind == "real" else 11))
    nbr = 300
    lens = rng.integers(0, 9, size=nbr)
    lens[rng.choice(nbr, 30, replace=False)] = 0
    lens[200:210] = nbr  # dense rows
    rows = np.repeat(np.arange(nbr), lens)
    cols = np.concatenate([rng.choice(nbr, k, replace=False) if k < nbr else np.arange(nbr) for k in lens])
    A = BSR.from_blocks(nbr, 1, rows, cols, rng.standard_normal((rows.size, 1, 1)))
    X = rng.standard_normal((nbr, R))
    if kind != "real":
        X = X + 1j * rng.standard_normal((nbr, R))
    gpu_ctx.upload_bsr(_lib.SLOT_H, A)
    y = gpu_ctx.bsr_mult(_lib.SLOT_H, X)
    ref = oracle.bsr_mult(A, X)
    assert rel_err(y, ref) < FP64_TOL
    assert np.array_equal(y[lens == 0], np.zeros_like(y[lens == 0]))


@pytest.mark.parametrize("b,L", [(1, 1), (1, 3), (1, 4), (1, 9), (2, 3)])
@pytest.mark.parametrize("R", [16, 33])
def test_uniform_rows(monkeypatch, oracle, b, L, R):
    # every block row with L blocks: the VALU SpMM computes the entry range
    # from the row index instead of loading it, which changes no sum (bit
    # identical to the loaded range); a ragged pattern uploaded into the same
    # buffers afterwards goes back to the loaded range
    rng = np.random.default_rng(100 * b + 10 * L + R)
    nbr = 257
    rows = np.repeat(np.arange(nbr), L)
    cols = np.concatenate([rng.choice(nbr, L, replace=False) for _ in range(nbr)])
    A = BSR.from_blocks(nbr, b, rows, cols, rng.standard_normal((rows.size, b, b)))
    X = rng.standard_normal((nbr * b, R)) + 1j * rng.standard_normal((nbr * b, R))
    lens = rng.multinomial(nbr * L, np.ones(nbr) / nbr)  # the same count, ragged (empty rows too)
    rows2 = np.repeat(np.arange(nbr), lens)
    cols2 = np.concatenate([rng.choice(nbr, k, replace=False) for k in lens])
    A2 = BSR.from_blocks(nbr, b, rows2, cols2, rng.standard_normal((rows2.size, b, b)))
    assert (lens != L).any()
    got = {}
    try:
        for on in ("1", "0"):
            monkeypatch.setenv("RG_ULEN", on)
            ctx = _lib.Context(0)  # the knob is read at context creation (process-wide)
            try:
                ctx.upload_bsr(_lib.SLOT_H, A)
                y = ctx.bsr_mult(_lib.SLOT_H, X)
                ctx.upload_bsr(_lib.SLOT_H, A2)
                y2 = ctx.bsr_mult(_lib.SLOT_H, X)
            finally:
                ctx.close()
            got[on] = (y, y2)
    finally:
        monkeypatch.setenv("RG_ULEN", "1")
        _lib.Context(0).close()
    assert rel_err(got["1"][0], oracle.bsr_mult(A, X)) < FP64_TOL
    assert rel_err(got["1"][1], oracle.bsr_mult(A2, X)) < FP64_TOL
    assert np.array_equal(got["1"][0], got["0"][0]) and np.array_equal(got["1"][1], got["0"][1])


@pytest.mark.parametrize("cplx", [False, True])
def test_fp32_value_upload(gpu_ctx, cplx):
    # F32 / C64 values are widened to FP64 exactly: the stored slot and the
    # products equal those of the widened values uploaded as F64 / C128
    rng = np.random.default_rng(41 + cplx)
    A = random_bsr(40, 3, rng, complex_=cplx)
    lo = A.values.astype(np.complex64 if cplx else np.float32)
    A32 = BSR(A.n_block_rows, A.block_dim, A.row_ptr, A.col_ind, lo)
    A64 = BSR(A.n_block_rows, A.block_dim, A.row_ptr, A.col_ind, lo.astype(np.complex128 if cplx else np.float64))
    X = rng.standard_normal((A.n_rows, 8)) + 1j * rng.standard_normal((A.n_rows, 8))
    gpu_ctx.upload_bsr(_lib.SLOT_H, A32)
    got = gpu_ctx.download_bsr(_lib.SLOT_H)
    y32 = gpu_ctx.bsr_mult(_lib.SLOT_H, X)
    gpu_ctx.upload_bsr(_lib.SLOT_H, A64)
    y64 = gpu_ctx.bsr_mult(_lib.SLOT_H, X)
    assert np.array_equal(got.values, A64.values) and np.array_equal(y32, y64)
    assert gpu_ctx.bsr_info(_lib.SLOT_H)["is_complex"] == cplx
