"""The precision form (FP32 vectors and matrices, FP32 MFMA/VALU products,
FP64 dots of exact FP32 products) against the FP64 path on the same inputs.

Tolerances are error budgets of the FP32 recurrence, stated per test: the
recurrence rounds every new vector to FP32 (eps_32 = 6e-8) and its error
grows with the step count, so per-moment errors are a few 1e-6 at a few
hundred moments. The DOS figure is max_E |rho32 - rho64| / max_E rho64,
the measure the precision target is stated in."""
import numpy as np
import pytest

from rsatb_amd import _lib
from rsatb_amd.lattice import diamond_supercell, graphene_rect_supercell, neighbor_edges
from rsatb_amd.models import RandomBlockModel, SkPzModel, assemble

pytestmark = pytest.mark.gpu


def dos_from(mu, npts=2001):
    x = -1.0 + np.arange(npts) / ((npts - 1) / 2.0)
    return _lib.kpm_density(mu, x, 0.5 / ((npts - 1) / 2.0))


def report(name, mu64, mu32):
    dmu = float(np.max(np.abs(mu32 - mu64)))
    r64, r32 = dos_from(mu64), dos_from(mu32)
    ddos = float(np.max(np.abs(r32 - r64)) / np.max(r64))
    print(f"\n{name}: max|dmu| = {dmu:.3e}, max|drho|/max rho = {ddos:.3e}", flush=True)
    return dmu, ddos


def test_fp32_graphene_orthogonal(gpu_ctx):
    st = graphene_rect_supercell(32, 32)
    H = assemble(st, neighbor_edges(st, 1.6), SkPzModel())
    gpu_ctx.upload_bsr(_lib.SLOT_H, H)
    gpu_ctx.set_operator(_lib.FORM_ORTHOGONAL, 0.0, 8.1 / 0.99)
    gpu_ctx.set_samples(st.n_atoms, 16, _lib.SAMPLES_RADEMACHER, 20240917)
    M = 512
    mu64 = gpu_ctx.kpm_moments(M, _lib.PREC_FP64)
    mu32 = gpu_ctx.kpm_moments(M, _lib.PREC_FP32)
    dmu, ddos = report("graphene 32x32 R=16 M=512", mu64, mu32)
    assert abs(mu32[0] - 1.0) < 1e-7  # <psi|psi> of FP32-rounded samples, summed in FP64
    assert dmu < 2e-5
    assert ddos < 1e-4


def test_fp32_diamond_sandwich(gpu_ctx):
    st = diamond_supercell(2)
    e = neighbor_edges(st, 2.7)
    m = RandomBlockModel()
    gpu_ctx.upload_bsr(_lib.SLOT_H, assemble(st, e, m, "H"))
    gpu_ctx.upload_bsr(_lib.SLOT_S, assemble(st, e, m, "S"))
    gpu_ctx.isqrt_patches()
    lo, hi = gpu_ctx.estimate_bounds(_lib.FORM_SANDWICH, steps=200, padding=0.5)
    gpu_ctx.set_operator(_lib.FORM_SANDWICH, (hi + lo) / 2, (hi - lo) / 2)
    gpu_ctx.set_samples(st.n_atoms * 13, 64, _lib.SAMPLES_PHASE, 20240917)
    M = 512
    mu64 = gpu_ctx.kpm_moments(M, _lib.PREC_FP64)
    mu32 = gpu_ctx.kpm_moments(M, _lib.PREC_FP32)
    dmu, ddos = report("diamond 2x2x2 sandwich R=64 M=512", mu64, mu32)
    assert dmu < 2e-5
    assert ddos < 1e-4


def test_fp32_matches_fp64_on_exact_fixture(gpu_ctx):
    # test_dos.cpp:69-139 fixture: the +-0.5 levels are exact in FP32 and the
    # +-1/10 samples round once; cos(n pi/3) is reached to FP32 accuracy
    n = 100
    from rsatb_amd.bsr import BSR

    H = BSR.from_blocks(n, 1, np.arange(n), np.arange(n),
                        np.array([0.5 if g % 2 == 0 else -0.5 for g in range(n)]).reshape(-1, 1, 1))
    psi = np.array([(1.0 if (g * 7 + 3) % 5 < 2 else -1.0) / 10.0 for g in range(n)])
    gpu_ctx.upload_bsr(_lib.SLOT_H, H)
    gpu_ctx.set_operator(_lib.FORM_ORTHOGONAL, 0.0, 1.0)
    gpu_ctx.set_samples(n, 1, _lib.SAMPLES_UPLOAD_REAL, data=psi)
    mu = gpu_ctx.kpm_moments(64, _lib.PREC_FP32)
    exact = np.array([np.cos(k * np.pi / 3.0) if k % 2 == 0 else 0.0 for k in range(64)])
    assert np.max(np.abs(mu - exact)) < 1e-5


@pytest.mark.parametrize("case", ["mfma_b13_R128", "valu_b1_R16"])
def test_fp32_dots_are_compensated(monkeypatch, case):
    # mu_0 of the precision form is <psi|psi> over the FP32-rounded samples:
    # a sum of products of FP32 values, each exact in FP64. With samples over
    # 2^-20 .. 2^20 and long per-lane sums (the MFMA case runs on 8
    # workgroups: about 1,700 terms per lane) a plain FP64 sum drifts by many
    # ulps (16.7 measured); the MFMA kernels' compensated sums (TwoSum per
    # term, per lane, then compensated folds and workgroup sums) stay within
    # a few ulps of the correctly rounded sum. The VALU kernel (b = 1 here)
    # sums a few rows per lane and compensates across workgroups only.
    import math

    rng = np.random.default_rng(5)
    if case == "mfma_b13_R128":
        monkeypatch.setenv("RG_MFMA_GRID", "8")
        st = diamond_supercell(12)  # 13,824 atoms x 13 orbitals, MFMA kernel with 16 tiles
        H = assemble(st, neighbor_edges(st, 2.7), RandomBlockModel(), "H")
        R = 128
    else:
        st = graphene_rect_supercell(64, 64)
        H = assemble(st, neighbor_edges(st, 1.6), SkPzModel())
        R = 16
    rows = H.n_rows
    scale = 2.0 ** rng.integers(-20, 21, size=(rows, R))
    psi = (rng.standard_normal((rows, R)) + 1j * rng.standard_normal((rows, R))) * scale
    ctx = _lib.Context(0)
    try:
        ctx.upload_bsr(_lib.SLOT_H, H)
        ctx.set_operator(_lib.FORM_ORTHOGONAL, 0.0, 64.0)
        ctx.set_samples(rows, R, _lib.SAMPLES_UPLOAD_COMPLEX, data=psi)
        _, cols = ctx.kpm_moments(4, _lib.PREC_FP32, per_column=True)
    finally:
        ctx.close()
        if case == "mfma_b13_R128":  # the process-wide grid cap back to its default
            monkeypatch.setenv("RG_MFMA_GRID", "0")
            _lib.Context(0).close()
    f = psi.astype(np.complex64)
    exact = np.array([math.fsum(np.concatenate([f[:, u].real.astype(np.float64) ** 2,
                                                f[:, u].imag.astype(np.float64) ** 2])) for u in range(R)])
    rel = np.abs(cols[0].real - exact) / exact
    print(f"\n{case}: max rel. error of mu_0 over {rows} rows: {rel.max():.2e} "
          f"({rel.max() / 2.0 ** -53:.1f} ulp)", flush=True)
    assert np.all(np.abs(cols[0].imag) == 0.0)
    assert rel.max() <= 4 * 2.0 ** -53
