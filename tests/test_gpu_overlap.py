"""The differentiable sandwich form with a learnable overlap: the
graph-local factor Z = S^{-1/2} built on the device (one-wave patch kernel for
patches up to 16 x 16), dL/dH, dL/dZ and dL/dS from
rg_kpm_moments_backward_ex against the dense FP64 eigendecomposition chain
(oracle.sandwich_moment_gradients: Daleckii-Krein for T_m and for x^{-1/2} on
every patch, each pinned by central differences on the CPU), dL/dtheta for
graphene with onsite, three hopping shells and the nearest-neighbour overlap
s1 against central differences of device forwards, the errors a stale forward
raises, and device memory across repeated forward/backward calls."""
import numpy as np
import pytest

from rsatb_amd import _lib
from rsatb_amd.autograd import (DifferentiableKPM, kpm_dos_torch, sk_pz_overlap_parameterization,
                                sk_pz_parameterization)
from rsatb_amd.bsr import BSR
from rsatb_amd.lattice import graphene_rect_supercell, neighbor_edges

pytestmark = pytest.mark.gpu

THETA_H = np.array([0.0, -2.7, -0.2, -0.18])
S1 = 0.1


def graphene_models(cells):
    st = graphene_rect_supercell(cells, cells)
    hm = sk_pz_parameterization(st, neighbor_edges(st, 2.9), [1.6, 2.5, 2.9])
    sm = sk_pz_overlap_parameterization(st, neighbor_edges(st, 1.6), [1.6])
    return st, hm, sm


def sandwich_window(oracle, H, S, pad=1.1):
    Z = oracle.isqrt_patches(S).to_dense()
    A = Z.T @ H.to_dense() @ Z
    r = np.max(np.abs(np.linalg.eigvalsh(0.5 * (A + A.T))))
    return (-pad * r, pad * r)


def sym_on_pattern(G_entries, P):
    D = BSR(P.n_block_rows, P.block_dim, P.row_ptr, P.col_ind, G_entries).to_dense()
    Sd = D + D.T
    b = P.block_dim
    return np.array([Sd[r * b:(r + 1) * b, c * b:(c + 1) * b] for r, c in zip(P.block_rows(), P.col_ind)])


def rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def test_small_patch_factor_matches_eigen_route(oracle, gpu_ctx):
    # the one-wave kernel (n = 4 here) against the clamped eigendecomposition
    st, hm, sm = graphene_models(6)
    S = sm.build([S1])
    gpu_ctx.upload_bsr(_lib.SLOT_S, S)
    it, res = gpu_ctx.isqrt_patches(1e-12, 60)
    print(f"\nsmall patches: {it} updates, residual {res:.2e}", flush=True)
    Z = gpu_ctx.download_bsr(_lib.SLOT_Z)
    ZH = gpu_ctx.download_bsr(_lib.SLOT_ZH)
    Zo = oracle.isqrt_patches(S)
    assert np.array_equal(Z.col_ind, Zo.col_ind)
    assert np.max(np.abs(Z.values - Zo.values)) < 1e-12
    assert np.max(np.abs(ZH.to_dense() - Z.to_dense().T)) == 0.0
    assert res < 1e-12 and 0 < it < 60


@pytest.mark.parametrize("nb,b", [(5, 3), (4, 4)])
def test_small_patch_block_sizes(oracle, gpu_ctx, nb, b):
    # block patterns of nb blocks per row, patch dimension nb * b <= 16
    # (4 x 4 = 16 fills the kernel's tile exactly)
    # a ring of 12 block rows: neighbours i +- 1, i +- 2 (5 blocks a row) or
    # i +- 1 and the opposite row i + 6 (4 blocks a row)
    rng = np.random.default_rng(nb * 10 + b)
    n = 12
    rows, cols, blks = [], [], []
    offs = [1, 2] if nb == 5 else [1]
    for i in range(n):
        B = rng.standard_normal((b, b))
        rows.append(i); cols.append(i); blks.append(np.eye(b) + 0.02 * (B + B.T))
    for i in range(n):
        for k in offs:
            j = (i + k) % n
            B = 0.05 * rng.standard_normal((b, b))
            rows += [i, j]; cols += [j, i]; blks += [B, B.T]
    if nb == 4:
        for i in range(n // 2):
            B = 0.05 * rng.standard_normal((b, b))
            rows += [i, i + n // 2]; cols += [i + n // 2, i]; blks += [B, B.T]
    S = BSR.from_blocks(n, b, np.array(rows), np.array(cols), np.array(blks))
    assert int(np.max(np.diff(S.row_ptr))) == nb and nb * b <= 16
    gpu_ctx.upload_bsr(_lib.SLOT_S, S)
    gpu_ctx.isqrt_patches(1e-12, 60)
    Z = gpu_ctx.download_bsr(_lib.SLOT_Z)
    assert np.max(np.abs(Z.values - oracle.isqrt_patches(S).values)) < 1e-12


def test_overlap_gradients_match_dense_chain_16x16(oracle):
    # the 16 x 16 rectangular cell (1024 atoms): every slot's gradient and
    # dL/dtheta against the dense FP64 eigendecomposition chain
    st, hm, sm = graphene_models(16)
    H, S = hm.build(THETA_H), sm.build([S1])
    window = sandwich_window(oracle, H, S)
    M, R = 128, 8
    eng = DifferentiableKPM(hm, R, window, M, checkpoint_stride=8, sample_kind="rademacher", overlap=sm)
    try:
        theta = np.concatenate([THETA_H, [S1]])
        mu = eng.forward(theta)
        psi = eng.ctx.get_samples()
        g = np.random.default_rng(5).standard_normal(M)
        outs = eng.ctx.kpm_moments_backward_ex(g, want_h=True, want_z=True, want_s=True)
        grad = eng.backward(g)
    finally:
        eng.ctx.close()
    c, a = (window[0] + window[1]) / 2, (window[1] - window[0]) / 2
    ref = oracle.sandwich_moment_gradients(H, S, c, a, psi, g)
    assert np.max(np.abs(mu - ref["mu"].real)) < 1e-11
    e_h = rel(sym_on_pattern(outs["H"], H), sym_on_pattern(ref["H"], H))
    e_z = rel(outs["Z"], ref["Z"])
    e_s = rel(outs["S"], ref["S"])
    ref_theta = np.concatenate([hm.contract(ref["H"]), sm.contract(ref["S"])])
    e_t = rel(grad, ref_theta)
    print(f"\nrelative errors: dH {e_h:.2e} dZ {e_z:.2e} dS {e_s:.2e} dtheta {e_t:.2e}; dL/ds1 = {grad[-1]:.6e}",
          flush=True)
    assert e_h <= 1e-10 and e_z <= 1e-10 and e_s <= 1e-10 and e_t <= 1e-10


def test_overlap_dos_loss_against_finite_differences(oracle):
    # the C4 loss (L2 between Jackson DOS and the DOS of parameters perturbed
    # by a seeded +-5 %) with s1 learnable: torch autograd against central
    # differences of device forwards (patches rebuilt on every forward)
    import torch

    st, hm, sm = graphene_models(8)
    theta0 = np.concatenate([THETA_H, [S1]])
    rng = np.random.default_rng(20240917)
    target_theta = theta0 * (1.0 + 0.05 * rng.choice([-1.0, 1.0], size=theta0.size))
    window = max((sandwich_window(oracle, hm.build(t[:4]), sm.build(t[4:]), 1.15) for t in (theta0, target_theta)),
                 key=lambda w: w[1])
    M = 128
    eng = DifferentiableKPM(hm, 16, window, M, checkpoint_stride=32, sample_kind="phase", overlap=sm)
    x = np.linspace(-0.95, 0.95, 301)
    try:
        with torch.no_grad():
            target = kpm_dos_torch(torch.from_numpy(eng.forward(target_theta)), x, 0.5 / 150)

        def loss_of(th):
            mu = torch.from_numpy(eng.forward(th))
            return float(torch.sum((kpm_dos_torch(mu, x, 0.5 / 150) - target) ** 2))

        th = torch.tensor(theta0, dtype=torch.float64, requires_grad=True)
        loss = torch.sum((kpm_dos_torch(eng.moments(th), x, 0.5 / 150) - target) ** 2)
        loss.backward()
        grad = th.grad.numpy().copy()
        h = 1e-5
        fd = np.zeros_like(grad)
        for k in range(theta0.size):
            e = np.zeros_like(theta0)
            e[k] = h
            fd[k] = (loss_of(theta0 + e) - loss_of(theta0 - e)) / (2 * h)
    finally:
        eng.ctx.close()
    print("\nautograd", grad, "\nfinite-diff", fd, flush=True)
    assert np.max(np.abs(grad - fd)) <= 1e-6 * np.max(np.abs(fd))
    assert abs(grad[-1]) > 1e-3 * np.max(np.abs(grad))


def test_two_forwards_in_one_graph():
    # a loss over two thetas through one engine: each node's backward uses
    # its own forward (re-run when a later forward replaced the states)
    import torch

    st, hm, sm = graphene_models(4)
    eng = DifferentiableKPM(hm, 4, (-12.0, 12.0), 32, checkpoint_stride=4, overlap=sm)
    try:
        t1 = np.concatenate([THETA_H, [S1]])
        t2 = t1 * 1.03
        w = np.linspace(1.0, -1.0, 32)
        x1 = torch.tensor(t1, requires_grad=True)
        x2 = torch.tensor(t2, requires_grad=True)
        loss = torch.sum(torch.from_numpy(w) * eng.moments(x1)) + 2.0 * torch.sum(
            torch.from_numpy(w) * eng.moments(x2))
        loss.backward()
        eng.forward(t1)
        g1 = eng.backward(w)
        eng.forward(t2)
        g2 = eng.backward(2.0 * w)
    finally:
        eng.ctx.close()
    assert np.max(np.abs(x1.grad.numpy() - g1)) <= 1e-12 * np.max(np.abs(g1))
    assert np.max(np.abs(x2.grad.numpy() - g2)) <= 1e-12 * np.max(np.abs(g2))


def test_stale_forward_states_are_rejected():
    st, hm, sm = graphene_models(4)
    ctx = _lib.Context(0)
    try:
        ctx.upload_bsr(_lib.SLOT_S, sm.build([S1]))
        ctx.isqrt_patches()
        H = hm.build(THETA_H)
        ctx.upload_bsr(_lib.SLOT_H, H)
        ctx.set_operator(_lib.FORM_SANDWICH, 0.0, 12.0)
        ctx.set_samples(H.n_rows, 4, _lib.SAMPLES_RADEMACHER, 3)
        ctx.set_checkpointing(4)
        g = np.ones(16)
        ctx.kpm_moments(16)
        ctx.kpm_moments_backward_ex(g, want_h=True, want_s=True)  # valid
        for change in ("H", "window", "samples", "S"):
            if change == "H":
                ctx.upload_bsr(_lib.SLOT_H, H)
            elif change == "window":
                ctx.set_operator(_lib.FORM_SANDWICH, 0.0, 12.5)
            elif change == "samples":
                ctx.set_samples(H.n_rows, 4, _lib.SAMPLES_RADEMACHER, 3)
            if change != "S":
                with pytest.raises(_lib.RsatbGpuError, match="stale"):
                    ctx.kpm_moments_backward(g)
                ctx.kpm_moments(16)
            else:
                # a new overlap without new patches: dL/dS has no factor to go through
                ctx.upload_bsr(_lib.SLOT_S, sm.build([0.11]))
                with pytest.raises(_lib.RsatbGpuError, match="rg_isqrt_patches"):
                    ctx.kpm_moments_backward_ex(g, want_h=True, want_s=True)
                ctx.kpm_moments_backward_ex(g, want_h=True)  # H, Z unchanged: still valid
        # a larger operator after the forward: stale, never a read past the states
        st2, hm2, sm2 = graphene_models(6)
        ctx.upload_bsr(_lib.SLOT_H, hm2.build(THETA_H))
        with pytest.raises(_lib.RsatbGpuError):
            ctx.kpm_moments_backward(g)
    finally:
        ctx.close()


def test_repeated_forward_backward_holds_device_memory():
    # 50 forward + backward calls (patches rebuilt each time) on a C4-shaped
    # graph: device free memory returns to its level after the first call
    st, hm, sm = graphene_models(32)
    eng = DifferentiableKPM(hm, 8, (-12.0, 12.0), 64, checkpoint_stride=8, overlap=sm)
    try:
        theta = np.concatenate([THETA_H, [S1]])
        w = np.linspace(1.0, -1.0, 64)
        eng.forward(theta)
        eng.backward(w)
        free0, _ = eng.ctx.mem_info()
        for _ in range(50):
            eng.forward(theta)
            eng.backward(w)
        free1, _ = eng.ctx.mem_info()
    finally:
        eng.ctx.close()
    print(f"\nfree memory change over 50 calls: {(free1 - free0) / 2**20:.3f} MiB", flush=True)
    assert abs(free1 - free0) <= 1 << 20
