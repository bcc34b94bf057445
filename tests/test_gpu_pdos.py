"""Projected DOS moments on the GPU (rg_kpm_pdos_moments) against the
oracle's kpm_pdos_moments on the same samples, against the dense
group-resolved recurrence on a complete basis (rsatb/calc/test/test_pdos.cpp),
and through size-independent properties at the diamond 8^3 size: the sum
rule over a partition, the all-rows channel equal to the DOS moments, and
mu_0^g of +-1 samples."""
import numpy as np
import pytest

from pdos_cases import K_HI, K_LO, N_ATOMS, NORB, as_bsr, chain
from rsatb_amd import _lib
from rsatb_amd.bsr import BSR
from rsatb_amd.driver import RSATB, SetupOptions
from rsatb_amd.lattice import diamond_supercell, graphene_rect_supercell, neighbor_edges
from rsatb_amd.models import RandomBlockModel, SkPzModel, assemble
from rsatb_amd.pdos import pdos_groups_by_atom, pdos_groups_by_orbital_slot

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12


def _sandwich_ctx(oracle, cells, R, seed=11):
    st = diamond_supercell(cells)
    e = neighbor_edges(st, 2.7)
    m = RandomBlockModel()
    H = assemble(st, e, m, "H")
    S = assemble(st, e, m, "S")
    ctx = _lib.Context(0)
    ctx.upload_bsr(_lib.SLOT_S, S)
    ctx.isqrt_patches()
    Z = ctx.download_bsr(_lib.SLOT_Z)
    ctx.upload_bsr(_lib.SLOT_H, H)
    ctx.set_operator(_lib.FORM_SANDWICH, 0.0, 30.0)
    ctx.set_samples(H.n_rows, R, _lib.SAMPLES_PHASE, seed)
    return ctx, st, H, Z


@pytest.mark.parametrize("grouping", ["atom", "slot"])
def test_pdos_complete_basis_matches_dense(oracle, grouping):
    """test_pdos.cpp:262-298 on the GPU: identity samples, sandwich form."""
    H, S = chain()
    Z = as_bsr(oracle.isqrt_patches(S))
    N = N_ATOMS * NORB
    M = 40
    gor = np.arange(N) // NORB if grouping == "atom" else np.arange(N) % NORB
    G = N_ATOMS if grouping == "atom" else NORB
    c, a = 0.5 * (K_HI + K_LO), 0.5 * (K_HI - K_LO)
    ctx = _lib.Context(0)
    try:
        ctx.upload_bsr(_lib.SLOT_H, H)
        ctx.upload_bsr(_lib.SLOT_Z, Z)
        ctx.upload_bsr(_lib.SLOT_ZH, Z.conj_transpose())
        ctx.set_operator(_lib.FORM_SANDWICH, c, a)
        ctx.set_samples(N, N, _lib.SAMPLES_UPLOAD_REAL, 0, np.eye(N))
        total, group = ctx.kpm_pdos_moments(M, gor, G)
    finally:
        ctx.close()
    A = oracle.dense_scaled(1, H, c, a, Z).real
    ref = oracle.dense_group_moments(A, gor, G, M)
    assert np.max(np.abs(N * group.real - ref[:G])) < 1e-8
    assert np.max(np.abs(N * total.real - ref[G])) < 1e-8
    assert np.max(np.abs(group.sum(axis=0) - total)) < 1e-12


@pytest.mark.parametrize("doubling", [True, False])
@pytest.mark.parametrize("M", [1, 2, 3, 4, 5, 64])
def test_pdos_orthogonal_matches_oracle(oracle, gpu_ctx, doubling, M):
    st = graphene_rect_supercell(8, 8)
    H = assemble(st, neighbor_edges(st, 1.6), SkPzModel(shells=((1.6, -2.7, 0.0),)), "H")
    N = H.n_rows
    gpu_ctx.upload_bsr(_lib.SLOT_H, H)
    gpu_ctx.set_operator(_lib.FORM_ORTHOGONAL, 0.0, 8.1 / 0.99)
    gpu_ctx.set_samples(N, 16, _lib.SAMPLES_RADEMACHER, 20240917)
    # ragged: every 7th row in no group, group 4 empty
    gor = np.where(np.arange(N) % 7 == 0, -1, np.arange(N) % 4)
    dbl = _lib.DOUBLING_ON if doubling else _lib.DOUBLING_OFF
    total, group = gpu_ctx.kpm_pdos_moments(M, gor, 5, _lib.PREC_FP64, dbl)
    op = oracle.Operator(0, H, 0.0, 8.1 / 0.99)
    psi = oracle.samples_rademacher(20240917, N, 16)
    ot, og = op.pdos_moments(psi, M, gor, 5, doubling=doubling)
    assert np.max(np.abs(total - ot)) <= FP64_TOL
    assert np.max(np.abs(group - og)) <= FP64_TOL
    assert np.all(group[4] == 0)
    mu = gpu_ctx.kpm_moments(M, _lib.PREC_FP64, dbl)
    assert np.array_equal(total, mu)  # the same loop and the same dots


@pytest.mark.parametrize("grouping", ["atom", "slot"])
def test_pdos_sandwich_matches_oracle(oracle, grouping):
    ctx, st, H, Z = _sandwich_ctx(oracle, 1, 32)
    try:
        at = np.zeros(st.n_atoms, dtype=int)
        g = pdos_groups_by_atom(at, [13]) if grouping == "atom" else pdos_groups_by_orbital_slot(at, [13])
        M = 48
        total, group = ctx.kpm_pdos_moments(M, g.group_of_row, g.n_groups)
        psi = ctx.get_samples(as_complex=True)
    finally:
        ctx.close()
    op = oracle.Operator(1, H, 0.0, 30.0, Z=Z)
    ot, og = op.pdos_moments(psi, M, g.group_of_row, g.n_groups)
    assert np.max(np.abs(total - ot)) <= FP64_TOL
    assert np.max(np.abs(group - og)) <= FP64_TOL
    assert np.max(np.abs(group.sum(axis=0) - total)) <= 1e-13


def test_pdos_fp32_against_fp64():
    """Precision form: per-group moments within the FP32 bound of the DOS
    moments (|dmu| < 2e-5)."""
    ctx, st, H, Z = _sandwich_ctx(None, 2, 64)
    try:
        g = pdos_groups_by_orbital_slot(np.zeros(st.n_atoms, dtype=int), [13])
        t64, g64 = ctx.kpm_pdos_moments(256, g.group_of_row, g.n_groups, _lib.PREC_FP64)
        t32, g32 = ctx.kpm_pdos_moments(256, g.group_of_row, g.n_groups, _lib.PREC_FP32)
    finally:
        ctx.close()
    assert np.max(np.abs(t32 - t64)) < 2e-5
    assert np.max(np.abs(g32 - g64)) < 2e-5
    assert np.max(np.abs(g32.sum(axis=0) - t32)) < 1e-12


def test_pdos_diamond8_properties():
    """Diamond 8^3 sandwich (53,248 rows, R = 64 complex): sum rule over the
    atom partition, total = the DOS moments bit for bit, and mu_0^g = 13/N
    for the per-atom groups of unit-modulus phase samples."""
    ctx, st, H, Z = _sandwich_ctx(None, 8, 64, seed=20240917)
    try:
        g = pdos_groups_by_atom(np.zeros(st.n_atoms, dtype=int), [13])
        M = 64
        total, group = ctx.kpm_pdos_moments(M, g.group_of_row, g.n_groups)
        mu = ctx.kpm_moments(M)
    finally:
        ctx.close()
    N = H.n_rows
    assert np.array_equal(total, mu)
    assert np.max(np.abs(group.sum(axis=0) - total)) < 1e-13
    assert np.max(np.abs(group[:, 0] - 13.0 / N)) < 1e-15


def test_pdos_errors(gpu_ctx):
    H = BSR.from_blocks(4, 1, np.arange(4), np.arange(4), np.ones((4, 1, 1)))
    gpu_ctx.upload_bsr(_lib.SLOT_H, H)
    gpu_ctx.set_operator(_lib.FORM_ORTHOGONAL, 0.0, 2.0)
    gpu_ctx.set_samples(4, 2, _lib.SAMPLES_RADEMACHER, 1)
    with pytest.raises(_lib.RsatbGpuError, match="exceeds n_groups"):
        gpu_ctx.kpm_pdos_moments(4, [0, 1, 2, 3], 3)
    with pytest.raises(_lib.RsatbGpuError, match="every scalar row"):
        gpu_ctx.kpm_pdos_moments(4, [0, 1, 2], 3)
    # no groups at all: the total alone
    t, g = gpu_ctx.kpm_pdos_moments(4, [-1, -1, -1, -1], 0)
    assert g.shape == (0, 4) and np.array_equal(t, gpu_ctx.kpm_moments(4))


def test_compute_pdos_driver():
    """RSATB.compute_pdos: groups sum to the total, the total is compute_dos's
    DOS on the same grid (observables.hpp:446-560)."""
    st = graphene_rect_supercell(16, 16)
    model = SkPzModel(shells=((1.6, -2.7, 0.0),))
    calc = RSATB(st, model, SetupOptions(nsamples=16, energy_bounds=(-8.6, 8.6)))
    out = calc.compute_pdos(0.1, "atom")
    dos, energy = calc.compute_dos(0.1)
    assert out["n_group"] == st.n_atoms and out["population"] == "lowdin"
    assert out["pdos"].shape == (1, st.n_atoms, energy.size)
    assert np.allclose(out["energy"], energy, rtol=0, atol=1e-12)
    assert np.max(np.abs(out["total"][0] - dos)) <= 1e-12 * np.max(np.abs(dos))
    assert np.max(np.abs(out["pdos"][0].sum(axis=0) - out["total"][0])) <= 1e-10 * np.max(np.abs(dos))
    slot = calc.compute_pdos(0.1, "orbital_slot")
    assert slot["n_group"] == 1 and list(slot["group_type"]) == [0] and list(slot["group_slot"]) == [0]
    with pytest.raises(RuntimeError, match="projection"):
        calc.compute_pdos(0.1, "shell")
