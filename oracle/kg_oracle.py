"""CPU restatement of the reference's Kubo-Greenwood DC spectrum, orthogonal
basis (TEST INFRASTRUCTURE ONLY: tests/ and the bench's CPU timing use it as
the checker; the product never imports it).

Follows rsatb/calc/fftcheb_transport.hpp (Greenwood formulation) and its
moment reference rsatb/calc/test/test_fftcheb_transport.cpp:156-271:
  - sequences (:290-430 of fftcheb_transport.hpp; test :156-230): with the
    scaled operator A = (H - c) / a and s = 1 / (2 a),
        R_m^x = T_m(A) (-s V_x r)       (current, ac.hpp:36-45)
        L_m^x = -s V_x T_m(A) r         (current_adjoint, ac.hpp:49-57)
    the orthogonal form applies V itself in both directions (cheby.hpp:258-262);
  - greenwood_moments: kernel(eps_k) = sum_{m,n} g_m T_m(eps_k) g_n T_n(eps_k)
    Re<L_m|R_n> / max(1e-14, 1 - eps_k^2), g_m = 2 j_m / (pi (1 + delta_m0)),
    eps_k = cos(pi (k + 1/2) / M), summed over samples and divided by their
    count, reported on the ascending grid times the prefactor
    (fftcheb_transport.hpp:570-623);
  - greenwood_fftcheb: the same quantity by the row-wise DCT-III route of
    accumulate_fftcheb_block (:436-466): y_k = x_0 + 2 sum_m x_m
    cos(pi m (k + 1/2) / M) of (j_m / pi) seq_m per row (fft.hpp:125-136);
  - bond_velocity: the orthogonal commutator velocity at Gamma,
    V_a = i sum_R (R + r_j - r_i)_a H_ij(R) (operator/commutator_velocity.hpp:
    7-43, AccumulateBondMoment with kDisplacement and factor i).
Pinned in tests/test_kg.py: the reference's toy problem (fftcheb route equal
to the moment reference, test_fftcheb_transport.cpp:511-521), the DCT-III
against the explicit cosine sum (:523-545), the complete-sample moments
against the dense spectral sum  sum_{n,n'} T_m(E_n) T_m'(E_n') |V_nn'|^2,
and the velocity's eigenbasis diagonal against -dE/dk (Hellmann-Feynman,
commutator_velocity.hpp:33-38).
"""
from __future__ import annotations

import numpy as np


def jackson(M: int) -> np.ndarray:
    """jackson_kernel (rsatb/calc/kpm.hpp:18-26)."""
    n = np.arange(M)
    q = np.pi / (M + 1)
    return ((M - n + 1) * np.cos(q * n) + np.sin(q * n) / np.tan(q)) / (M + 1)


def nodes(M: int) -> np.ndarray:
    return np.cos(np.pi * (np.arange(M) + 0.5) / M)


def sequences(H, V, r, center: float, half: float, M: int):
    """(L, R): [3][M] arrays of the sample shape (r: [rows] or [rows, S])."""
    try:
        import scipy.sparse as sp
    except ImportError:  # dense only
        sp = None
    if sp is not None and sp.issparse(H):
        A = (H - center * sp.identity(H.shape[0], format="csr")) / half
    else:
        A = (np.asarray(H) - center * np.eye(H.shape[0])) / half
    s = 1.0 / (2.0 * half)
    r = np.asarray(r, dtype=complex)
    L = [[None] * M for _ in range(3)]
    R = [[None] * M for _ in range(3)]
    b_prev = b_cur = None
    r_prev = r_cur = None
    for m in range(M):
        if m == 0:
            b = r
            rr = [-s * (V[a] @ r) for a in range(3)]
        elif m == 1:
            b = A @ b_cur
            rr = [A @ x for x in r_cur]
        else:
            b = 2.0 * (A @ b_cur) - b_prev
            rr = [2.0 * (A @ x) - y for x, y in zip(r_cur, r_prev)]
        b_prev, b_cur = b_cur, b
        r_prev, r_cur = r_cur, rr
        for a in range(3):
            L[a][m] = -s * (V[a] @ b)
            R[a][m] = rr[a]
    return L, R


def _dot(x, y) -> complex:
    return complex(np.sum(np.conj(x) * y))


def moment_matrix(L, R, M: int) -> np.ndarray:
    """mu[a][m, n] = <L_m^a | R_n^a> (summed over sample columns)."""
    return np.array([[[_dot(L[a][m], R[a][n]) for n in range(M)] for m in range(M)] for a in range(3)])


def greenwood_from_moments(mu: np.ndarray, M: int, n_samples: int = 1, prefactor: float = 1.0,
                           kernel: np.ndarray | None = None) -> np.ndarray:
    """sigma[k, a] on the ascending grid from mu[a] (complex M x M)."""
    j = jackson(M) if kernel is None else kernel
    g = 2.0 * j / (np.pi * (1.0 + (np.arange(M) == 0)))
    eps = nodes(M)
    T = np.cos(np.outer(np.arange(M), np.arccos(eps)))  # T[m, k]
    W = g[:, None] * T                                   # W[m, k]
    out = np.zeros((M, 3))
    for a in range(3):
        native = np.einsum("mk,mn,nk->k", W, mu[a].real, W) / np.maximum(1e-14, 1.0 - eps * eps)
        out[:, a] = prefactor * native[::-1] / n_samples
    return out


def dct_type3(x: np.ndarray) -> np.ndarray:
    """y_k = x_0 + 2 sum_{m>=1} x_m cos(pi m (k + 1/2) / n), via an FFT of the
    even extension (fft.hpp:125-158)."""
    x = np.asarray(x, dtype=complex)
    n = x.shape[-1]
    ext = np.zeros(x.shape[:-1] + (2 * n,), dtype=complex)
    ext[..., 0] = x[..., 0]
    m = np.arange(1, n)
    th = np.pi * m / (2.0 * n)
    ext[..., m] = x[..., 1:] * np.exp(1j * th)
    ext[..., 2 * n - m] = x[..., 1:] * np.exp(-1j * th)
    return (np.fft.ifft(ext, axis=-1) * (2 * n))[..., :n]


def greenwood_fftcheb(L, R, M: int, n_samples: int = 1, prefactor: float = 1.0) -> np.ndarray:
    """Row-wise DCT-III route (fftcheb_transport.hpp:436-466, 570-623)."""
    j = jackson(M)
    eps = nodes(M)
    out = np.zeros((M, 3))
    for a in range(3):
        Ls = np.stack([np.asarray(L[a][m]) for m in range(M)], axis=-1)  # [rows(, S), M]
        Rs = np.stack([np.asarray(R[a][m]) for m in range(M)], axis=-1)
        pl = dct_type3(Ls * (j / np.pi))
        pr = dct_type3(Rs * (j / np.pi))
        acc = np.sum((np.conj(pl) * pr).reshape(-1, M), axis=0)
        native = acc / np.maximum(1e-14, 1.0 - eps * eps)
        out[:, a] = prefactor * native.real[::-1] / n_samples
    return out


def energies(M: int, center: float, half: float) -> np.ndarray:
    """Ascending physical energies of the nodes (fftcheb_transport.hpp:616-622)."""
    return center + half * nodes(M)[::-1]


def bond_velocity(n_atoms: int, b: int, edges_src, edges_dst, bond_vecs, edge_blocks):
    """Dense V_a = i sum over edges of bond_a H_edge (Gamma, orthogonal basis);
    edge_blocks[e] is the row-major b x b block of edge e."""
    V = np.zeros((3, n_atoms * b, n_atoms * b), dtype=complex)
    for e in range(len(edges_src)):
        i, j = int(edges_src[e]), int(edges_dst[e])
        for a in range(3):
            V[a, i * b:(i + 1) * b, j * b:(j + 1) * b] += 1j * bond_vecs[e][a] * edge_blocks[e]
    return V


def bloch_h_and_v(n_atoms: int, b: int, onsite_blocks, edges_src, edges_dst, bond_vecs, edge_blocks, k_cart):
    """H(k) and V_a(k) = i sum (bond)_a H_edge e^{-i k.bond} in the atomic
    gauge (commutator_velocity.hpp:24-31, the phase 2 pi k.(R + s_ji) written
    with the Cartesian bond vector)."""
    k = np.asarray(k_cart, dtype=np.float64)
    H = np.zeros((n_atoms * b, n_atoms * b), dtype=complex)
    V = np.zeros((3, n_atoms * b, n_atoms * b), dtype=complex)
    for a in range(n_atoms):
        H[a * b:(a + 1) * b, a * b:(a + 1) * b] += onsite_blocks[a]
    for e in range(len(edges_src)):
        i, j = int(edges_src[e]), int(edges_dst[e])
        d = np.asarray(bond_vecs[e], dtype=np.float64)
        ph = np.exp(-1j * float(k @ d))
        H[i * b:(i + 1) * b, j * b:(j + 1) * b] += edge_blocks[e] * ph
        for ax in range(3):
            V[ax, i * b:(i + 1) * b, j * b:(j + 1) * b] += 1j * d[ax] * edge_blocks[e] * ph
    return H, V


PAIRS = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]  # (xx, xy, xz, yy, yz, zz)


def moment_pairs(L, R, M: int) -> np.ndarray:
    """mu[q][m, n] = <L_m^a | R_n^b> for the six (a, b) of PAIRS."""
    return np.array([[[_dot(L[a][m], R[b][n]) for n in range(M)] for m in range(M)] for a, b in PAIRS])


def _integrate(kern_native: np.ndarray, M: int, prefactor: float) -> np.ndarray:
    """Ascending grid, cumulative trapezoid of Re(i kernel) (fftcheb_transport.hpp:627-652)."""
    eps = nodes(M)[::-1]
    out = np.zeros((M, 6))
    running = np.zeros(6)
    prev = np.zeros(6)
    for k in range(M):
        integrand = (1j * kern_native[M - 1 - k]).real
        if k > 0:
            running += 0.5 * (prev + integrand) * (eps[k] - eps[k - 1])
        out[k] = prefactor * running
        prev = integrand
    return out


def bastin_from_moments(mu6: np.ndarray, M: int, n_samples: int = 1, prefactor: float = 1.0) -> np.ndarray:
    """The moment reference of test_fftcheb_transport.cpp:271-361:
    phi1 = sum_m g_m e^{i m theta} seq_m, phi2 = sum_m m g_m e^{i m theta} seq_m,
    phi3 = sum_m g_m cos(m theta) seq_m, kernel = eps (<L3|R1> + <L1|R3>)
    + i sin(theta) (<L2|R3> - <L3|R2>), then the cumulative integral."""
    j = jackson(M)
    g = 2.0 * j / (np.pi * (1.0 + (np.arange(M) == 0)))
    m = np.arange(M)
    kern = np.zeros((M, 6), dtype=complex)
    for k in range(M):
        th = np.pi * (k + 0.5) / M
        c = g * np.cos(m * th)
        e = g * np.exp(1j * m * th)
        em = m * e
        for q in range(6):
            mu = mu6[q]
            l3r1 = c @ mu @ e
            l1r3 = np.conj(e) @ mu @ c
            l2r3 = np.conj(em) @ mu @ c
            l3r2 = c @ mu @ em
            kern[k, q] = np.cos(th) * (l3r1 + l1r3) + 1j * np.sin(th) * (l2r3 - l3r2)
    return _integrate(kern / n_samples, M, prefactor)


def bastin_fftcheb(L, R, M: int, n_samples: int = 1, prefactor: float = 1.0) -> np.ndarray:
    """Row-wise route of accumulate_fftcheb_block, Bastin (:436-540): phi3 by
    the DCT-III, phi1 and phi2 by zero-padded inverse FFTs of length 2M."""
    j = jackson(M)
    g = 2.0 * j / (np.pi * (1.0 + (np.arange(M) == 0)))
    m = np.arange(M)
    pw = g * np.exp(1j * np.pi * m / (2.0 * M))
    eps = nodes(M)
    th = np.pi * (np.arange(M) + 0.5) / M

    def zpifft(x):
        pad = np.zeros(x.shape[:-1] + (2 * M,), dtype=complex)
        pad[..., :M] = x
        return (np.fft.ifft(pad, axis=-1) * (2 * M))[..., :M]

    phi = {}
    for side, seqs in (("L", L), ("R", R)):
        for a in range(3):
            S = np.stack([np.asarray(seqs[a][mm]) for mm in range(M)], axis=-1)
            phi[side, a, 3] = dct_type3(S * (j / np.pi))
            phi[side, a, 1] = zpifft(S * pw)
            phi[side, a, 2] = zpifft(S * (pw * m))
    kern = np.zeros((M, 6), dtype=complex)
    for q, (a, b) in enumerate(PAIRS):
        def dot(x, y):
            return np.sum((np.conj(x) * y).reshape(-1, M), axis=0)
        kern[:, q] = eps * (dot(phi["L", a, 3], phi["R", b, 1]) + dot(phi["L", a, 1], phi["R", b, 3])) + \
            1j * np.sin(th) * (dot(phi["L", a, 2], phi["R", b, 3]) - dot(phi["L", a, 3], phi["R", b, 2]))
    return _integrate(kern / n_samples, M, prefactor)


def dc_temperature_quadrature(spectrum, energy, mu: float, temp: float, per_segment: int = 4000):
    """sigma(mu, T) = int -df/dE sigma(E) dE over the grid's span with sigma
    linear between the grid energies (integrate_dc_spectrum,
    rsatb/calc/fftcheb_transport.hpp:978-1040), by composite Simpson on a
    fine subdivision of every segment with the analytic
    -df/dE = f (1 - f) / kT, kT = 8.617333262e-5 temp eV. An independent route
    to the closed form the reference uses; temp must be positive."""
    e = np.asarray(energy, dtype=np.float64)
    s = np.asarray(spectrum, dtype=np.float64).reshape(len(e), -1)
    kt = 8.617333262e-5 * temp
    n = per_segment + (per_segment % 2)
    out = np.zeros(s.shape[1])
    for i in range(1, len(e)):
        x = np.linspace(e[i - 1], e[i], n + 1)
        z = (x - mu) / kt
        q = np.exp(-np.abs(z))
        w = q / (kt * (1.0 + q) ** 2)  # -df/dE = f (1 - f) / kT
        t = (x - e[i - 1]) / (e[i] - e[i - 1])
        y = w[:, None] * ((1 - t)[:, None] * s[i - 1] + t[:, None] * s[i])
        h = (e[i] - e[i - 1]) / n
        out += h / 3 * (y[0] + y[-1] + 4 * y[1:-1:2].sum(axis=0) + 2 * y[2:-1:2].sum(axis=0))
    return out
