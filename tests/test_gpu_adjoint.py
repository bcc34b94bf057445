"""The adjoint recurrence (rg_kpm_moments_backward) against the dense FP64
Daleckii-Krein derivative of the same stochastic moments, against central
finite differences through the forward pass, and across checkpoint strides.

Per-entry gradients of the doubled recurrence are compared in their
symmetric part (G_ij + G_ji), which is all a symmetric Hamiltonian sees."""
import numpy as np
import pytest

from rsatb_amd import _lib
from rsatb_amd.autograd import DifferentiableKPM, kpm_dos_torch, sk_pz_parameterization
from rsatb_amd.bsr import BSR
from rsatb_amd.lattice import diamond_supercell, graphene_rect_supercell, neighbor_edges
from rsatb_amd.models import RandomBlockModel, assemble

pytestmark = pytest.mark.gpu


def sym_on_pattern(G_entries, H):
    """(G + G^T) at every stored entry, from per-entry gradients."""
    D = BSR(H.n_block_rows, H.block_dim, H.row_ptr, H.col_ind, G_entries).to_dense()
    S = D + D.T
    b = H.block_dim
    rows = H.block_rows()
    return np.array([S[r * b:(r + 1) * b, c * b:(c + 1) * b] for r, c in zip(rows, H.col_ind)])


def dense_sym(G_dense, H):
    S = G_dense + G_dense.T
    b = H.block_dim
    rows = H.block_rows()
    return np.array([S[r * b:(r + 1) * b, c * b:(c + 1) * b] for r, c in zip(rows, H.col_ind)])


@pytest.mark.parametrize("kind", ["rademacher", "phase"])
def test_adjoint_orthogonal_matches_dense(oracle, kind):
    st = graphene_rect_supercell(4, 4)
    model = sk_pz_parameterization(st, neighbor_edges(st, 2.9), [1.6, 2.5, 2.9])
    theta = np.array([0.3, -2.7, -0.2, -0.18])
    H = model.build(theta)
    a = 1.2 * np.max(np.abs(np.linalg.eigvalsh(H.to_dense())))
    M = 64
    eng = DifferentiableKPM(model, 6, (-a, a), M, checkpoint_stride=8, sample_kind=kind)
    mu = eng.forward(theta)
    psi = eng.ctx.get_samples()
    g = np.random.default_rng(1).standard_normal(M)
    G = eng.ctx.kpm_moments_backward(g)
    Gd = oracle.dense_moment_gradient(H.to_dense() / a, psi, g) / a
    ref = dense_sym(Gd, H)
    got = sym_on_pattern(G, H)
    assert np.max(np.abs(got - ref)) <= 1e-10 * np.max(np.abs(ref))
    # and the forward moments themselves
    assert np.max(np.abs(mu - oracle.Operator(0, H, 0.0, a).moments(psi, M).real)) < 1e-12


def test_adjoint_sandwich_matches_dense(oracle):
    st = diamond_supercell(1)
    e = neighbor_edges(st, 2.7)
    m = RandomBlockModel()
    H = assemble(st, e, m, "H")
    S = assemble(st, e, m, "S")
    Z = oracle.isqrt_patches(S)
    Zb = BSR(Z.n_block_rows, Z.block_dim, Z.row_ptr, Z.col_ind, Z.values)
    A = oracle.dense_scaled(1, H, 0.0, 1.0, Z).real
    ev = np.linalg.eigvalsh(0.5 * (A + A.T))
    lo, hi = ev[0] - 0.5, ev[-1] + 0.5
    c, a = (hi + lo) / 2, (hi - lo) / 2
    ctx = _lib.Context(0)
    try:
        ctx.upload_bsr(_lib.SLOT_H, H)
        ctx.upload_bsr(_lib.SLOT_Z, Zb)
        ctx.upload_bsr(_lib.SLOT_ZH, Zb.conj_transpose())
        ctx.set_operator(_lib.FORM_SANDWICH, c, a)
        ctx.set_samples(H.n_rows, 8, _lib.SAMPLES_PHASE, 7)
        ctx.set_checkpointing(5)
        M = 40
        ctx.kpm_moments(M)
        psi = ctx.get_samples()
        g = np.random.default_rng(2).standard_normal(M)
        G = ctx.kpm_moments_backward(g)
    finally:
        ctx.close()
    At = (A - c * np.eye(A.shape[0])) / a
    GA = oracle.dense_moment_gradient(At, psi, g) / a
    Zd = Zb.to_dense()
    GH = Zd @ GA @ Zd.T  # dL/dH = Z (dL/dA) Z^T
    ref = dense_sym(GH, H)
    got = sym_on_pattern(G, H)
    assert np.max(np.abs(got - ref)) <= 1e-10 * np.max(np.abs(ref))


def test_checkpoint_stride_independence():
    st = graphene_rect_supercell(6, 6)
    model = sk_pz_parameterization(st, neighbor_edges(st, 1.6), [1.6])
    theta = np.array([0.0, -2.7])
    grads = []
    for K in (1, 3, 8, 32, 100):
        eng = DifferentiableKPM(model, 4, (-8.2, 8.2), 90, checkpoint_stride=K)
        eng.forward(theta)
        grads.append(eng.ctx.kpm_moments_backward(np.linspace(1.0, -1.0, 90)))
        eng.ctx.close()
    for G in grads[1:]:
        assert np.max(np.abs(G - grads[0])) <= 1e-13 * np.max(np.abs(grads[0]))


def test_autograd_dos_loss_against_finite_differences():
    import torch

    st = graphene_rect_supercell(8, 8)
    model = sk_pz_parameterization(st, neighbor_edges(st, 2.9), [1.6, 2.5, 2.9])
    theta0 = np.array([0.0, -2.7, -0.2, -0.18])
    rng = np.random.default_rng(20240917)
    target_theta = theta0 * (1.0 + 0.05 * rng.choice([-1.0, 1.0], size=theta0.size))
    target_theta[0] = 0.05
    bound = max(np.max(np.abs(np.linalg.eigvalsh(model.build(t).to_dense()))) for t in (theta0, target_theta))
    window = (-1.1 * bound, 1.1 * bound)
    M = 128
    eng = DifferentiableKPM(model, 16, window, M, checkpoint_stride=32, sample_kind="phase")
    x = np.linspace(-0.95, 0.95, 301)
    with torch.no_grad():
        target = kpm_dos_torch(torch.from_numpy(eng.forward(target_theta)), x, 0.5 / 150)

    def loss_of(theta_np):
        mu = torch.from_numpy(eng.forward(theta_np))
        return float(torch.sum((kpm_dos_torch(mu, x, 0.5 / 150) - target) ** 2))

    th = torch.tensor(theta0, dtype=torch.float64, requires_grad=True)
    mu = eng.moments(th)
    loss = torch.sum((kpm_dos_torch(mu, x, 0.5 / 150) - target) ** 2)
    loss.backward()
    grad = th.grad.numpy().copy()
    h = 1e-5
    fd = np.zeros_like(grad)
    for k in range(theta0.size):
        e = np.zeros_like(theta0)
        e[k] = h
        fd[k] = (loss_of(theta0 + e) - loss_of(theta0 - e)) / (2 * h)
    print("\nautograd", grad, "\nfinite-diff", fd, flush=True)
    assert np.max(np.abs(grad - fd)) <= 1e-6 * np.max(np.abs(fd))
    eng.ctx.close()


def test_backward_without_forward_states_is_an_error():
    ctx = _lib.Context(0)
    try:
        H = BSR.from_blocks(4, 1, np.arange(4), np.arange(4), np.ones((4, 1, 1)))
        ctx.upload_bsr(_lib.SLOT_H, H)
        ctx.set_operator(_lib.FORM_ORTHOGONAL, 0.0, 2.0)
        ctx.set_samples(4, 2, _lib.SAMPLES_RADEMACHER, 1)
        ctx.kpm_moments(8)
        with pytest.raises(_lib.RsatbGpuError, match="checkpointing"):
            ctx.kpm_moments_backward(np.ones(8))
    finally:
        ctx.close()
