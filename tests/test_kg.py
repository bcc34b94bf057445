"""Kubo-Greenwood oracle (CPU): the reference's FFTCheb Greenwood route and
its moment form, the DCT-III, the complete-sample moments against the dense
spectral sum, and the bond-moment velocity against Hellmann-Feynman."""
from __future__ import annotations

import math
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))

from oracle import kg_oracle as kg  # noqa: E402
from oracle import oracle as O  # noqa: E402
from oracle import sk_oracle as so  # noqa: E402
from sk_cases import periodic_edges, silicon_cell, sp3_silicon_table  # noqa: E402


def toy_greenwood():
    """make_greenwood_problem of test_fftcheb_transport.cpp:77-113."""
    H = np.array([[-0.35, 0.18 - 0.07j], [0.18 + 0.07j, 0.42]])
    Vx = np.array([[0.0, 1.0], [1.0, 0.0]], dtype=complex)
    Vy = np.array([[0.2, 0.1j], [-0.1j, -0.15]])
    Vz = np.array([[0.1, 0.0], [0.0, -0.1]], dtype=complex)
    a = np.array([1 / math.sqrt(2), 1j / math.sqrt(2)])
    b = np.array([0.3 - 0.4j, -0.2 + 0.5j])
    return H, [Vx, Vy, Vz], a, b


def test_jackson_matches_pinned_kernel():
    for M in (1, 2, 12, 64):
        assert np.allclose(kg.jackson(M), O.jackson_kernel(M), atol=1e-15)


def test_toy_fftcheb_equals_moment_reference():
    """test_fftcheb_transport.cpp:511-521 (window [-1, 1], M = 12)."""
    H, V, a, b = toy_greenwood()
    M = 12
    L, R = kg.sequences(H, V, a, 0.0, 1.0, M)
    mom = kg.greenwood_from_moments(kg.moment_matrix(L, R, M), M)
    fft = kg.greenwood_fftcheb(L, R, M)
    assert np.max(np.abs(mom - fft)) < 1e-13 * max(1.0, np.max(np.abs(mom)))
    # two samples: the average of single-sample spectra (:656-673)
    Lb, Rb = kg.sequences(H, V, b, 0.0, 1.0, M)
    two_L, two_R = kg.sequences(H, V, np.stack([a, b], axis=1), 0.0, 1.0, M)
    both = kg.greenwood_fftcheb(two_L, two_R, M, n_samples=2)
    avg = 0.5 * (fft + kg.greenwood_fftcheb(Lb, Rb, M))
    assert np.max(np.abs(both - avg)) < 1e-13


def test_dct_type3_explicit_sum():
    """test_fftcheb_transport.cpp:523-545."""
    n, batch = 8, 2
    x = np.array([[0.1 * (1 + b) * (m + 1) - 0.05j * (2 + b) * (m + 1) for m in range(n)] for b in range(batch)])
    y = kg.dct_type3(x)
    k = np.arange(n)
    ref = np.array([[x[b, 0] + 2 * sum(x[b, m] * math.cos(math.pi * m * (kk + 0.5) / n) for m in range(1, n))
                     for kk in k] for b in range(batch)])
    assert np.max(np.abs(y - ref)) < 1e-12


def test_complete_samples_equal_dense_spectral_sum():
    """With the unit vectors as samples, mu_mn = s^2 sum_pq T_m(e_p) T_n(e_q) |V_pq|^2
    in the eigenbasis of A = (H - c) / a."""
    rng = np.random.default_rng(3)
    N, M, c = 7, 9, 0.2
    X = rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N))
    H = (X + X.conj().T) / 2
    half = 1.05 * np.max(np.abs(np.linalg.eigvalsh(H) - c))  # spectrum inside the window
    V = []
    for _ in range(3):
        Y = rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N))
        V.append((Y + Y.conj().T) / 2)
    L, R = kg.sequences(H, V, np.eye(N, dtype=complex), c, half, M)
    mu = kg.moment_matrix(L, R, M)
    e, U = np.linalg.eigh((H - c * np.eye(N)) / half)
    T = np.cos(np.outer(np.arange(M), np.arccos(np.clip(e, -1, 1))))  # T[m, p]
    s2 = 1.0 / (4 * half * half)
    for a in range(3):
        Ve = U.conj().T @ V[a] @ U
        ref = s2 * T @ (np.abs(Ve) ** 2) @ T.T
        assert np.max(np.abs(mu[a] - ref)) < 1e-12 * max(1.0, np.max(np.abs(ref)))


def test_bond_velocity_hellmann_feynman():
    """V_a(k) = -dH/dk_a and its eigenbasis diagonal = -dE_n/dk_a
    (commutator_velocity.hpp:33-38) for the periodic spd Si cell."""
    t = sp3_silicon_table()
    pos, cell = silicon_cell()
    ei, sh = periodic_edges(pos, cell, 3.2)
    vecs = so.edge_vectors(pos, cell, ei, sh)
    blocks = [so.pair_block(t, "H", 0, 0, v) for v in vecs]
    blocks = [np.zeros((9, 9)) if bk is None else bk for bk in blocks]
    onsite = [np.diag(t.atom_types[0].h_onsite)] * len(pos)
    k = np.array([0.11, -0.07, 0.05])
    H, V = kg.bloch_h_and_v(len(pos), 9, onsite, ei[0], ei[1], vecs, blocks, k)
    assert np.allclose(H, H.conj().T, atol=1e-12)
    h = 1e-5
    for a in range(3):
        dk = np.zeros(3)
        dk[a] = h
        Hp, _ = kg.bloch_h_and_v(len(pos), 9, onsite, ei[0], ei[1], vecs, blocks, k + dk)
        Hm, _ = kg.bloch_h_and_v(len(pos), 9, onsite, ei[0], ei[1], vecs, blocks, k - dk)
        assert np.max(np.abs(V[a] + (Hp - Hm) / (2 * h))) < 1e-6
        E, U = np.linalg.eigh(H)
        Ep = np.linalg.eigvalsh(Hp)
        Em = np.linalg.eigvalsh(Hm)
        diag = np.real(np.einsum("pi,pq,qi->i", U.conj(), V[a], U))
        gap = np.min(np.diff(E))
        if gap > 1e-3:  # non-degenerate: the band derivative is defined
            assert np.max(np.abs(diag + (Ep - Em) / (2 * h))) < 1e-5
    # at Gamma the velocity is i times the real bond moment: Hermitian, imaginary
    _, V0 = kg.bloch_h_and_v(len(pos), 9, onsite, ei[0], ei[1], vecs, blocks, np.zeros(3))
    Vd = kg.bond_velocity(len(pos), 9, ei[0], ei[1], vecs, blocks)
    assert np.allclose(V0, Vd, atol=1e-13)
    assert np.allclose(V0.real, 0.0) and np.allclose(V0, V0.conj().transpose(0, 2, 1), atol=1e-12)


def toy_bastin():
    """make_bastin_problem of test_fftcheb_transport.cpp:115-152."""
    H = np.array([[0.05, 0.8j], [-0.8j, -0.1]])
    Vx = np.array([[0.0, 0.9 + 0.2j], [0.9 - 0.2j, 0.0]])
    Vy = np.array([[0.4, -0.7j], [0.7j, -0.3]])
    Vz = np.array([[0.2, 0.5], [0.5, -0.25]], dtype=complex)
    a = np.array([0.6, 0.8j])
    b = np.array([-0.5 + 0.1j, 0.2 + 0.7j])
    return H, [Vx, Vy, Vz], a, b


def test_toy_bastin_fftcheb_equals_moment_reference():
    """test_fftcheb_transport.cpp:547-557 (M = 12, window [-1, 1])."""
    H, V, a, b = toy_bastin()
    M = 12
    for r, ns in ((a, 1), (np.stack([a, b], axis=1), 2)):
        L, R = kg.sequences(H, V, r, 0.0, 1.0, M)
        mom = kg.bastin_from_moments(kg.moment_pairs(L, R, M), M, n_samples=ns, prefactor=3.0)
        fft = kg.bastin_fftcheb(L, R, M, n_samples=ns, prefactor=3.0)
        assert np.max(np.abs(mom - fft)) < 1e-12 * max(1.0, np.max(np.abs(mom)))
        assert np.all(mom[0] == 0.0)  # the integral starts at the lowest node


def test_linear_transport_model_velocity_basis():
    """Every parameter's velocity factor B_{a,k} equals the oracle's bond
    velocity V_a = i sum_e bond_a H_e of that parameter's edge entries
    (commutator_velocity.hpp:7-43), and H_k its hopping and onsite entries;
    H(theta), B_a(theta) are linear in theta."""
    from transport_cases import THETA, graphene_transport

    st, e, model, ec, oc = graphene_transport(3, 2)
    n = st.n_atoms
    for k in range(ec.shape[1]):
        ek = np.zeros(ec.shape[1])
        ek[k] = 1.0
        V = kg.bond_velocity(n, 1, e.src, e.dst, e.vec, [np.array([[c]]) for c in ec[:, k]])
        for a in range(3):
            np.testing.assert_allclose(1j * model.velocity[a].build(ek).to_dense(), V[a], rtol=0, atol=1e-13)
        Hd = np.diag(oc[:, k]).astype(float)
        np.add.at(Hd, (e.src, e.dst), ec[:, k])
        np.testing.assert_allclose(model.hamiltonian.build(ek).to_dense(), Hd, rtol=0, atol=1e-13)
    H = model.hamiltonian.build(THETA).to_dense()
    assert np.allclose(H, H.T)
    Bx = model.velocity[0].build(THETA).to_dense()
    assert np.allclose(Bx, -Bx.T)  # V = i B Hermitian
