"""Device Kubo-Greenwood (rg_sk_velocity + rg_kubo_greenwood) against the
oracle: velocity factors, the sample-summed moment matrices mu = <L_m|R_n>
and the Greenwood spectrum (both the moment form and the reference's
FFTCheb route), to 1e-10 relative (FP64; the device contracts mu with a
GEMM in another summation order)."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import kg_oracle as kg
from oracle import sk_oracle as so
from rsatb_amd import _lib
from rsatb_amd.bsr import BSR
from sk_cases import periodic_edges, silicon_cell, sp3_silicon_table

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = _lib.Context(0)
    yield c
    c.close()


def setup_si(ctx):
    t = sp3_silicon_table()
    pos, cell = silicon_cell()
    ei, sh = periodic_edges(pos, cell, 3.2)
    types = np.zeros(len(pos), dtype=np.int32)
    a = t.arrays("H")
    ctx.sk_assemble(_lib.SLOT_H, a, types, pos, cell, ei.T.copy(), sh)
    ctx.sk_velocity(a, types, pos, cell, ei.T.copy(), sh)
    rp, ci, blk = so.assemble(t, "H", types, pos, cell, ei, sh)
    H = BSR(len(pos), 9, rp, ci, blk).to_dense()
    vecs = so.edge_vectors(pos, cell, ei, sh)
    blocks = [so.pair_block(t, "H", 0, 0, v) for v in vecs]
    blocks = [np.zeros((9, 9)) if bk is None else bk for bk in blocks]
    V = kg.bond_velocity(len(pos), 9, ei[0], ei[1], vecs, blocks)
    return H, V


def test_velocity_factors(ctx):
    H, V = setup_si(ctx)
    for a, slot in enumerate((_lib.SLOT_VX, _lib.SLOT_VY, _lib.SLOT_VZ)):
        B = ctx.download_bsr(slot).to_dense()
        assert np.max(np.abs(1j * B - V[a])) < 1e-12


@pytest.mark.parametrize("kind", [_lib.SAMPLES_RADEMACHER, _lib.SAMPLES_PHASE])
def test_greenwood_against_oracle(ctx, kind):
    H, V = setup_si(ctx)
    e = np.linalg.eigvalsh(H)
    c, half = 0.5 * (e[0] + e[-1]), 0.55 * (e[-1] - e[0])
    ctx.set_operator(_lib.FORM_ORTHOGONAL, c, half)
    R, M = 4, 48
    ctx.set_samples(H.shape[0], R, kind, 11)
    print("case: Si spd cell KG, samples", kind, flush=True)
    sigma, energy, mu = ctx.kubo_greenwood(M, prefactor=2.5, with_moments=True)
    psi = ctx.get_samples()
    L, Rs = kg.sequences(H, V, psi, c, half, M)
    ref_mu = kg.moment_matrix(L, Rs, M)
    assert np.max(np.abs(mu - ref_mu)) < 1e-10 * np.max(np.abs(ref_mu))
    ref = kg.greenwood_from_moments(ref_mu, M, n_samples=R, prefactor=2.5)
    fft = kg.greenwood_fftcheb(L, Rs, M, n_samples=R, prefactor=2.5)
    scale = np.max(np.abs(ref))
    assert np.max(np.abs(sigma - ref)) < 1e-10 * scale
    assert np.max(np.abs(sigma - fft)) < 1e-10 * scale
    assert np.allclose(energy, kg.energies(M, c, half), atol=1e-12)


def test_greenwood_needs_velocity_and_orthogonal_form(ctx):
    c2 = _lib.Context(0)
    try:
        t = sp3_silicon_table()
        pos, cell = silicon_cell()
        ei, sh = periodic_edges(pos, cell, 3.2)
        c2.sk_assemble(_lib.SLOT_H, t.arrays("H"), np.zeros(len(pos), np.int32), pos, cell, ei.T.copy(), sh)
        c2.set_operator(_lib.FORM_ORTHOGONAL, 0.0, 30.0)
        c2.set_samples(72, 2, _lib.SAMPLES_RADEMACHER, 1)
        with pytest.raises(_lib.RsatbGpuError, match="velocity"):
            c2.kubo_greenwood(8)
    finally:
        c2.close()


def test_driver_dc_spectrum():
    """RSATB.compute_dc_spectrum over an orthogonal SktbModel: the device
    spectrum equals the oracle's with the reference's prefactor
    pi g_s 4 N / volume and the S/m conversion (observables.hpp:722-733,
    rsatb/driver.py:390-396)."""
    from rsatb_amd.driver import RSATB, SetupOptions
    from rsatb_amd.lattice import Structure
    from rsatb_amd.sktb import SktbModel
    from sk_cases import silicon_params

    basis, h, _ = silicon_params()
    model = SktbModel(basis, h)
    pos, cell = silicon_cell()
    calc = RSATB(Structure(pos, cell), model, SetupOptions(nsamples=3, compute_bounds=True,
                                                          sample_kind="rademacher"))
    M = 40
    sigma, energy = calc.compute_dc_spectrum(M)
    assert sigma.shape == (1, M, 6)
    assert np.all(sigma[0][:, [1, 2, 4]] == 0.0)
    e = calc.edges
    ei = np.stack([e.src, e.dst])
    rp, ci, blk = so.assemble(model.table, "H", np.zeros(len(pos), np.int32), pos, cell, ei, e.shift)
    H = BSR(len(pos), 9, rp, ci, blk).to_dense()
    blocks = [so.pair_block(model.table, "H", 0, 0, v) for v in e.vec]
    blocks = [np.zeros((9, 9)) if bk is None else bk for bk in blocks]
    V = kg.bond_velocity(len(pos), 9, e.src, e.dst, e.vec, blocks)
    emin, emax = calc.energy_bounds()
    c, half = (emax + emin) / 2, (emax - emin) / 2
    psi = calc.context.get_samples()
    L, Rs = kg.sequences(H, V, psi, c, half, M)
    pref = np.pi * 2 * 4 * H.shape[0] / abs(np.linalg.det(cell)) * 2.4341348e6
    ref = kg.greenwood_fftcheb(L, Rs, M, n_samples=3, prefactor=pref)
    assert np.max(np.abs(sigma[0][:, [0, 3, 5]] - ref)) < 1e-10 * np.max(np.abs(ref))
    assert np.allclose(energy, kg.energies(M, c, half), atol=1e-12)
    with pytest.raises(RuntimeError, match="formulation"):
        calc.compute_dc_spectrum(M, formulation="kubo")


def test_driver_dc_temperature():
    """RSATB.compute_dc: the device spectrum integrated against -df/dE at
    (mu, T) equals the oracle's spectrum integrated by quadrature
    (observables.hpp:792-854, fftcheb_transport.hpp:978-1040); the windowed
    "qe" route keeps the nodes near mu and agrees at a temperature whose
    window holds every node that -df/dE weighs; qe rejects Bastin."""
    from rsatb_amd.driver import RSATB, SetupOptions
    from rsatb_amd.lattice import Structure
    from rsatb_amd.sktb import SktbModel
    from sk_cases import silicon_params

    basis, h, _ = silicon_params()
    model = SktbModel(basis, h)
    pos, cell = silicon_cell()
    calc = RSATB(Structure(pos, cell), model, SetupOptions(nsamples=2, compute_bounds=True,
                                                          sample_kind="rademacher"))
    M = 48
    sigma_spec, energy = calc.compute_dc_spectrum(M)
    mu = energy[20] + 0.6 * (energy[21] - energy[20])
    for temp in (300.0, 3000.0):
        dc = calc.compute_dc(M, mu, temp)
        assert dc.shape == (1, 6)
        ref = kg.dc_temperature_quadrature(sigma_spec[0], energy, mu, temp, per_segment=2000)
        assert np.max(np.abs(dc[0] - ref)) <= 1e-9 * np.max(np.abs(sigma_spec))
        qe = calc.compute_dc(M, mu, temp, method="qe")
        assert np.max(np.abs(qe[0] - dc[0])) <= 1e-6 * np.max(np.abs(sigma_spec))
    dc0 = calc.compute_dc(M, mu, 0.0)
    t = (mu - energy[20]) / (energy[21] - energy[20])
    np.testing.assert_allclose(dc0[0], (1 - t) * sigma_spec[0, 20] + t * sigma_spec[0, 21], rtol=1e-12, atol=0)
    with pytest.raises(RuntimeError, match="greenwood"):
        calc.compute_dc(M, mu, 300.0, formulation="bastin", method="qe")


@pytest.mark.parametrize("kind", [_lib.SAMPLES_RADEMACHER, _lib.SAMPLES_PHASE])
def test_bastin_against_oracle(ctx, kind):
    """rg_dc_spectrum, Bastin: the six cross-axis moment matrices and the
    integrated spectrum against the oracle's moment form and FFTCheb route."""
    H, V = setup_si(ctx)
    e = np.linalg.eigvalsh(H)
    c, half = 0.5 * (e[0] + e[-1]), 0.55 * (e[-1] - e[0])
    ctx.set_operator(_lib.FORM_ORTHOGONAL, c, half)
    R, M = 3, 40
    ctx.set_samples(H.shape[0], R, kind, 5)
    print("case: Si spd cell Bastin, samples", kind, flush=True)
    sigma, energy, mu = ctx.dc_spectrum(M, formulation=1, prefactor=1.7, with_moments=True)
    psi = ctx.get_samples()
    L, Rs = kg.sequences(H, V, psi, c, half, M)
    ref_mu = kg.moment_pairs(L, Rs, M)
    assert np.max(np.abs(mu - ref_mu)) < 1e-10 * np.max(np.abs(ref_mu))
    ref = kg.bastin_from_moments(ref_mu, M, n_samples=R, prefactor=1.7)
    fft = kg.bastin_fftcheb(L, Rs, M, n_samples=R, prefactor=1.7)
    scale = np.max(np.abs(ref))
    assert np.max(np.abs(sigma - ref)) < 1e-10 * scale
    assert np.max(np.abs(sigma - fft)) < 1e-10 * scale
    # Greenwood through the same entry: the diagonal of the 6-component layout
    sg, _ = ctx.dc_spectrum(M, formulation=0, prefactor=1.7)
    g3 = kg.greenwood_from_moments(kg.moment_matrix(L, Rs, M), M, n_samples=R, prefactor=1.7)
    assert np.max(np.abs(sg[:, [0, 3, 5]] - g3)) < 1e-10 * np.max(np.abs(g3))
    assert np.all(sg[:, [1, 2, 4]] == 0.0)


def _oracle_spectrum(model, theta, psi, c, half, M, formulation, n_samples):
    H = model.hamiltonian.build(theta).to_dense()
    V = np.stack([1j * model.velocity[a].build(theta).to_dense() for a in range(3)])
    L, Rs = kg.sequences(H, V, psi, c, half, M)
    if formulation == "greenwood":
        return kg.greenwood_fftcheb(L, Rs, M, n_samples=n_samples)
    return kg.bastin_fftcheb(L, Rs, M, n_samples=n_samples)


@pytest.mark.parametrize("formulation", ["greenwood", "bastin"])
def test_dc_tangent_against_oracle_finite_differences(formulation):
    """rg_dc_spectrum_tangent: the Jacobian d sigma / d theta of a graphene
    p_z spectrum (onsite and three hopping shells, velocity from the same
    parameters) equals central differences of the oracle's spectrum
    (h = 1e-5, fixed window) to 1e-6 of its largest entry, and the spectrum
    equals the oracle's to 1e-10."""
    from rsatb_amd.autograd import DifferentiableConductivity
    from transport_cases import THETA, graphene_transport

    st, e, model, ec, oc = graphene_transport(4, 4)
    H = model.hamiltonian.build(THETA).to_dense()
    a = 1.3 * np.max(np.abs(np.linalg.eigvalsh(H)))
    M, S = 40, 3
    eng = DifferentiableConductivity(model, S, (-a, a), M, formulation=formulation)
    sigma, jac = eng.jacobian(THETA)
    psi = eng.ctx.get_samples()
    comps = [0, 3, 5] if formulation == "greenwood" else list(range(6))
    ref = _oracle_spectrum(model, THETA, psi, 0.0, a, M, formulation, S)
    assert np.max(np.abs(sigma[:, comps] - ref)) <= 1e-10 * np.max(np.abs(ref))
    h = 1e-5
    for k in range(THETA.size):
        tp, tm = THETA.copy(), THETA.copy()
        tp[k] += h
        tm[k] -= h
        fd = (_oracle_spectrum(model, tp, psi, 0.0, a, M, formulation, S)
              - _oracle_spectrum(model, tm, psi, 0.0, a, M, formulation, S)) / (2 * h)
        scale = max(np.max(np.abs(fd)), 1e-30)
        assert np.max(np.abs(jac[k][:, comps] - fd)) <= 1e-6 * scale, (k, np.max(np.abs(jac[k][:, comps] - fd)), scale)
    # every call recomputes the same primal spectrum
    np.testing.assert_allclose(eng.spectrum(THETA), sigma, rtol=0, atol=1e-12 * np.max(np.abs(sigma)))


def test_dc_conductivity_torch_gradient():
    """DifferentiableConductivity.conductivity under torch autograd: the
    gradient of L = sum_k w_k sigma_xx(E_k) + sigma_yy(E_mid)^2 equals
    central differences of L through the device spectrum."""
    import torch

    from rsatb_amd.autograd import DifferentiableConductivity
    from transport_cases import THETA, graphene_transport

    st, e, model, ec, oc = graphene_transport(4, 4)
    H = model.hamiltonian.build(THETA).to_dense()
    a = 1.3 * np.max(np.abs(np.linalg.eigvalsh(H)))
    M = 32
    eng = DifferentiableConductivity(model, 2, (-a, a), M, sample_kind="phase")
    w = torch.from_numpy(np.random.default_rng(3).standard_normal(M))

    def loss_of(sig):
        return (w * sig[:, 0]).sum() + sig[M // 2, 3] ** 2

    theta = torch.tensor(THETA, requires_grad=True)
    loss = loss_of(eng.conductivity(theta))
    loss.backward()
    g = theta.grad.numpy()
    h = 1e-5
    fd = np.zeros_like(THETA)
    for k in range(THETA.size):
        tp, tm = THETA.copy(), THETA.copy()
        tp[k] += h
        tm[k] -= h
        fd[k] = (loss_of(torch.from_numpy(eng.spectrum(tp))) - loss_of(torch.from_numpy(eng.spectrum(tm)))).item() / (2 * h)
    assert np.max(np.abs(g - fd)) <= 1e-6 * np.max(np.abs(fd))


def test_dc_tangent_needs_tangent_slots():
    """The tangent entry point rejects missing tangent matrices."""
    from transport_cases import THETA, graphene_transport

    st, e, model, ec, oc = graphene_transport(2, 2)
    c2 = _lib.Context(0)
    c2.set_samples(st.n_atoms, 1, _lib.SAMPLES_RADEMACHER, 1)
    c2.set_operator(_lib.FORM_ORTHOGONAL, 0.0, 12.0)
    c2.upload_bsr(_lib.SLOT_H, model.hamiltonian.build(THETA))
    for ax in range(3):
        c2.upload_bsr(_lib.SLOT_VX + ax, model.velocity[ax].build(THETA))
    with pytest.raises(RuntimeError, match="tangent"):
        c2.dc_spectrum_tangent(8)


@pytest.mark.parametrize("formulation", [0, 1])
@pytest.mark.parametrize("kind", [_lib.SAMPLES_RADEMACHER, _lib.SAMPLES_PHASE])
def test_fp32_transport_against_fp64(formulation, kind, monkeypatch):
    """rg_dc_spectrum_ex with RG_PREC_FP32 (FP32 recurrence and velocity
    products, FP32 stored sequences, FP64 moment GEMMs over widened
    K-chunks) against the FP64 spectrum of the same inputs: within 1e-5 of
    the largest FP64 value. With a small widening budget (many K-chunks,
    RG_KG_WIDEN_BYTES) the FP32 spectrum equals the one-chunk FP32 spectrum
    to 1e-12: the chunked FP64 accumulation only reorders the sums."""
    spectra = {}
    for budget in (None, "4096"):
        if budget is None:
            monkeypatch.delenv("RG_KG_WIDEN_BYTES", raising=False)
        else:
            monkeypatch.setenv("RG_KG_WIDEN_BYTES", budget)
        c = _lib.Context(0)
        try:
            H, V = setup_si(c)
            e = np.linalg.eigvalsh(H)
            c.set_operator(_lib.FORM_ORTHOGONAL, 0.5 * (e[0] + e[-1]), 0.55 * (e[-1] - e[0]))
            c.set_samples(H.shape[0], 3, kind, 5)
            print("case: Si KG fp32 vs fp64, formulation", formulation, "kind", kind, "budget", budget, flush=True)
            s64, _ = c.dc_spectrum(40, formulation, prefactor=1.0)
            s32, en = c.dc_spectrum(40, formulation, prefactor=1.0, precision=_lib.PREC_FP32)
            spectra[budget] = s32
        finally:
            c.close()
        scale = np.max(np.abs(s64))
        assert np.max(np.abs(s32 - s64)) <= 1e-5 * scale
        assert np.max(np.abs(s32 - s64)) > 0.0  # the FP32 path ran
    scale = np.max(np.abs(spectra[None]))
    assert np.max(np.abs(spectra["4096"] - spectra[None])) <= 1e-12 * scale


@pytest.mark.parametrize("formulation", [0, 1])
def test_dc_spectrum_at_k_against_atomic_gauge_oracle(formulation):
    """At k != 0 the device builds H(k) and B_a(k) with the lattice Bloch
    phase e^{2 pi i k . shift} per edge. The oracle's atomic-gauge pair
    (bloch_h_and_v, commutator_velocity.hpp:24-31, pinned by Hellmann-Feynman)
    at k' = -2 pi cell^-1 k is the same pair conjugated by U = diag(e^{i
    k_cart . r}); with the samples taken to U^H r the oracle's spectrum must
    equal the device's (1e-10 relative), for Greenwood and Bastin."""
    t = sp3_silicon_table()
    pos, cell = silicon_cell()
    ei, sh = periodic_edges(pos, cell, 3.2)
    types = np.zeros(len(pos), dtype=np.int32)
    a = t.arrays("H")
    kfrac = np.array([0.13, -0.21, 0.07])
    vecs = so.edge_vectors(pos, cell, ei, sh)
    blocks = [so.pair_block(t, "H", 0, 0, v) for v in vecs]
    blocks = [np.zeros((9, 9)) if bk is None else bk for bk in blocks]
    onsite = [np.diag(np.asarray(a["onsite"])[:9]) for _ in range(len(pos))]
    k_cart = 2 * np.pi * np.linalg.solve(cell, kfrac)
    H, V = kg.bloch_h_and_v(len(pos), 9, onsite, ei[0], ei[1], vecs, blocks, -k_cart)
    e = np.linalg.eigvalsh(H)
    c0, half = 0.5 * (e[0] + e[-1]), 0.55 * (e[-1] - e[0])
    M, R = 40, 3
    c = _lib.Context(0)
    try:
        c.sk_assemble(_lib.SLOT_H, a, types, pos, cell, ei.T.copy(), sh, kpoint=kfrac)
        c.sk_velocity(a, types, pos, cell, ei.T.copy(), sh, kpoint=kfrac)
        assert c.bsr_info(_lib.SLOT_VX)["is_complex"]
        c.set_operator(_lib.FORM_ORTHOGONAL, c0, half)
        c.set_samples(H.shape[0], R, _lib.SAMPLES_PHASE, 9)
        print("case: Si KG at k", kfrac, "formulation", formulation, flush=True)
        sigma, energy = c.dc_spectrum(M, formulation, prefactor=1.0)
        psi = c.get_samples()
    finally:
        c.close()
    ph = np.repeat(np.exp(-1j * (pos @ k_cart)), 9)  # U^H on every orbital row
    L, Rs = kg.sequences(H, V, ph[:, None] * psi, c0, half, M)
    if formulation == 0:
        ref = kg.greenwood_fftcheb(L, Rs, M, n_samples=R)
        got = sigma[:, [0, 3, 5]]
    else:
        ref = kg.bastin_fftcheb(L, Rs, M, n_samples=R)
        got = sigma
    scale = np.max(np.abs(ref))
    assert np.max(np.abs(got - ref)) <= 1e-10 * scale


def test_driver_dc_spectrum_kpoints():
    """RSATB.compute_dc_spectrum over two k-points: the Gamma entry equals a
    Gamma-only run, and the k != 0 entry is the device spectrum of H(k) with
    its own-gauge velocity (finite, and different from Gamma's)."""
    from rsatb_amd.driver import RSATB, SetupOptions
    from rsatb_amd.lattice import Structure
    from rsatb_amd.sktb import SktbModel
    from sk_cases import silicon_params

    basis, h, _ = silicon_params()
    model = SktbModel(basis, h)
    pos, cell = silicon_cell()
    opts = dict(nsamples=2, compute_bounds=False, energy_bounds=(-20.0, 25.0), sample_kind="phase")
    both = RSATB(Structure(pos, cell), model, SetupOptions(kpoints=[[0, 0, 0], [0.13, -0.21, 0.07]], **opts))
    gamma = RSATB(Structure(pos, cell), model, SetupOptions(**opts))
    s2, e2 = both.compute_dc_spectrum(32)
    s1, e1 = gamma.compute_dc_spectrum(32)
    assert s2.shape == (2, 32, 6)
    np.testing.assert_allclose(s2[0], s1[0], rtol=0, atol=1e-12 * np.max(np.abs(s1)))
    assert np.all(np.isfinite(s2[1]))
    assert np.max(np.abs(s2[1] - s2[0])) > 1e-6 * np.max(np.abs(s2[0]))
