"""The benchmark configurations themselves under -m gpu:

- plain DOS, graphene 64 x 64 rectangular cells (16,384 atoms), R = 16
  Rademacher samples (seed 20240917), M = 1024, 2001-point Jackson DOS,
  against the oracle (rsatb/calc/dos.hpp:124-189, kpm.hpp:206-241) at the
  exact size;
- the FP32 form at full depth: diamond 8^3 sandwich (53,248 orbitals), R = 64
  phase samples, M = 2048, DOS error of FP32 against FP64 <= 1e-5;
- twisted bilayer graphene at 1.05 degrees, Kubo-Greenwood (Greenwood) with
  the derivative along the interlayer hopping scale: one (31, 32) cell
  (11,908 atoms) against the oracle's sequences and central differences of
  the oracle's spectrum, and the 4 x 4 tile (190,528 atoms, R = 16,
  M = 512) FP32 against FP64 and the tangent against Richardson-extrapolated
  differences of device spectra."""
import numpy as np
import pytest

from rsatb_amd import _lib

pytestmark = pytest.mark.gpu

X2001 = -1.0 + np.arange(2001) / 1000.0


def dos_rel(mu_ref, mu):
    r_ref = _lib.kpm_density(mu_ref, X2001, 0.5 / 1000.0)
    r = _lib.kpm_density(mu, X2001, 0.5 / 1000.0)
    return float(np.max(np.abs(r - r_ref)) / np.max(r_ref))


def test_c1_graphene_64x64_exact_size(oracle):
    from rsatb_amd.lattice import graphene_rect_supercell, neighbor_edges
    from rsatb_amd.models import SkPzModel, assemble

    st = graphene_rect_supercell(64, 64)
    H = assemble(st, neighbor_edges(st, 1.6), SkPzModel())
    N, R, M, a = st.n_atoms, 16, 1024, 8.1 / 0.99
    print(f"\nC1: graphene 64x64, N={N}, R={R}, M={M}", flush=True)
    with _lib.Context(0) as c:
        c.upload_bsr(_lib.SLOT_H, H)
        c.set_operator(_lib.FORM_ORTHOGONAL, 0.0, a)
        c.set_samples(N, R, _lib.SAMPLES_RADEMACHER, 20240917)
        mu = c.kpm_moments(M)
        psi = c.get_samples(as_complex=False)
    psi_ref = oracle.samples_rademacher(20240917, N, R)
    assert np.array_equal(psi, psi_ref.real if np.iscomplexobj(psi_ref) else psi_ref)
    ref = oracle.Operator(oracle.FORM_ORTHOGONAL, H, 0.0, a).moments(psi_ref, M)
    assert np.max(np.abs(mu - ref)) <= 1e-10 * np.max(np.abs(ref))
    rho = _lib.kpm_density(mu, X2001, 0.5 / 1000.0)
    rho_ref = oracle.kpm_density(ref, X2001, 0.5 / 1000.0)
    assert np.max(np.abs(rho - rho_ref)) <= 1e-10 * np.max(np.abs(rho_ref))
    assert abs(mu[0].real - 1.0) < 1e-12


def test_fp32_dos_error_at_full_depth_c2():
    from rsatb_amd.driver import RSATB, PreconditionerOptions, SetupOptions
    from rsatb_amd.lattice import diamond_supercell
    from rsatb_amd.models import RandomBlockModel

    st = diamond_supercell(8)
    calc = RSATB(st, RandomBlockModel(), SetupOptions(
        nsamples=64, sample_seed=20240917, compute_bounds=True, sample_kind="phase",
        preconditioner=PreconditionerOptions(operator_form="sandwich")))
    calc._bind_kpoint(0)
    ctx = calc.context
    print("\nC2: diamond 8^3 sandwich, R=64, M=2048, FP32 vs FP64", flush=True)
    mu64 = ctx.kpm_moments(2048, _lib.PREC_FP64)
    mu32 = ctx.kpm_moments(2048, _lib.PREC_FP32)
    ctx.close()
    err = dos_rel(mu64, mu32)
    dmu = float(np.max(np.abs(mu32 - mu64)))
    print(f"DOS error {err:.3e}, max |dmu| {dmu:.3e}", flush=True)
    assert err <= 1e-5
    assert 0.0 < dmu < 1e-6  # the FP32 path ran, and tracks FP64 per moment


def _tblg_device(c, model, theta, e1, R, M, half, pref=1.0):
    c.set_samples(model.hamiltonian.pattern.n_rows, R, _lib.SAMPLES_RADEMACHER, 20240917)
    c.set_operator(_lib.FORM_ORTHOGONAL, 0.0, half)
    c.upload_bsr(_lib.SLOT_H, model.hamiltonian.build(theta))
    for a in range(3):
        c.upload_bsr(_lib.SLOT_VX + a, model.velocity[a].build(theta))
    c.upload_bsr(_lib.SLOT_DH, model.hamiltonian.build(e1))
    for a in range(3):
        c.upload_bsr(_lib.SLOT_DVX + a, model.velocity[a].build(e1))


def _gershgorin(model, theta):
    H = model.hamiltonian.build(theta)
    rows = np.repeat(np.arange(H.n_block_rows), np.diff(H.row_ptr))
    return float(np.max(np.bincount(rows, weights=np.abs(H.values.reshape(-1)), minlength=H.n_block_rows))) + 0.05


def test_tblg_cell_greenwood_and_interlayer_tangent_against_oracle():
    import scipy.sparse as sp

    from oracle import kg_oracle as kg
    from rsatb_amd.autograd import tblg_transport_model

    st, e, model = tblg_transport_model(1)
    N = st.n_atoms
    theta = np.array([1.0, 1.0])
    e1 = np.array([0.0, 1.0])
    R, M = 2, 32
    half = _gershgorin(model, theta)
    print(f"\nC5 cell: tBLG 1.05 deg, N={N}, R={R}, M={M}", flush=True)
    with _lib.Context(0) as c:
        _tblg_device(c, model, theta, e1, R, M, half)
        sigma, dsigma, energy = c.dc_spectrum_tangent(M, 0, _lib.KERNEL_JACKSON, 1.0)
        psi = c.get_samples()

    def csr(b, scale=1.0):
        return sp.csr_matrix((scale * b.values.reshape(-1), b.col_ind, b.row_ptr), shape=(N, N))

    def oracle_sigma(th):
        Hs = csr(model.hamiltonian.build(th))
        Vs = [csr(model.velocity[a].build(th), 1j) for a in range(3)]
        L, Rs = kg.sequences(Hs, Vs, psi, 0.0, half, M)
        return kg.greenwood_from_moments(kg.moment_matrix(L, Rs, M), M, n_samples=R)

    ref = oracle_sigma(theta)
    comps = [0, 3, 5]
    scale = np.max(np.abs(ref))
    assert np.max(np.abs(sigma[:, comps] - ref)) <= 1e-10 * scale
    h = 1e-5
    fd = (oracle_sigma(theta + h * e1) - oracle_sigma(theta - h * e1)) / (2 * h)
    err = np.max(np.abs(dsigma[:, comps] - fd)) / np.max(np.abs(fd))
    print(f"spectrum vs oracle {np.max(np.abs(sigma[:, comps] - ref)) / scale:.2e}; tangent vs oracle FD {err:.2e}",
          flush=True)
    assert err <= 1e-6


def test_tblg_4x4_fp32_and_tangent_at_scale():
    from rsatb_amd.autograd import tblg_transport_model
    from rsatb_amd.transport import interpolate_spectrum_component

    st, e, model = tblg_transport_model(4)
    theta = np.array([1.0, 1.0])
    e1 = np.array([0.0, 1.0])
    R, M = 16, 512
    half = _gershgorin(model, theta)
    print(f"\nC5: tBLG 1.05 deg 4x4, N={st.n_atoms}, R={R}, M={M}", flush=True)
    with _lib.Context(0) as c:
        _tblg_device(c, model, theta, e1, R, M, half)
        s64, energy = c.dc_spectrum(M, 0, _lib.KERNEL_JACKSON, 1.0)
        s32, _ = c.dc_spectrum(M, 0, _lib.KERNEL_JACKSON, 1.0, precision=_lib.PREC_FP32)
        sigma, dsigma, _ = c.dc_spectrum_tangent(M, 0, _lib.KERNEL_JACKSON, 1.0)

        def bind(th):
            c.upload_bsr(_lib.SLOT_H, model.hamiltonian.build(th))
            for a in range(3):
                c.upload_bsr(_lib.SLOT_VX + a, model.velocity[a].build(th))

        def central(h):
            bind(theta + h * e1)
            sp_, _ = c.dc_spectrum(M, 0, _lib.KERNEL_JACKSON, 1.0)
            bind(theta - h * e1)
            sm_, _ = c.dc_spectrum(M, 0, _lib.KERNEL_JACKSON, 1.0)
            return (sp_ - sm_) / (2 * h)

        fd = (4.0 * central(1e-4) - central(2e-4)) / 3.0
    comps = [0, 3]
    e32 = np.max(np.abs(s32 - s64)) / np.max(np.abs(s64))
    et = np.max(np.abs(dsigma[:, comps] - fd[:, comps])) / np.max(np.abs(fd[:, comps]))
    d0 = interpolate_spectrum_component(energy, dsigma, 0.0)[comps]
    f0 = interpolate_spectrum_component(energy, fd, 0.0)[comps]
    e0 = np.max(np.abs(d0 - f0)) / np.max(np.abs(f0))
    print(f"FP32 vs FP64 {e32:.2e}; tangent vs Richardson FD {et:.2e} (spectrum), {e0:.2e} (E_F = 0)", flush=True)
    assert 0.0 < e32 <= 1e-5
    assert np.max(np.abs(sigma - s64)) <= 1e-10 * np.max(np.abs(s64))
    assert et <= 1e-6 and e0 <= 1e-6
