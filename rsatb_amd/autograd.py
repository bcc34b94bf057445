"""Differentiable KPM moments: a torch.autograd.Function over the library's
adjoint Chebyshev recurrence (rg_kpm_moments_backward).

PyTorch is used only here, as the autograd wrapper: the forward and the
backward both run in the gfx950 library; torch sees the moments (a real
vector) and hands back dL/dmu. Parameters enter through a linear
parameterization of the Hamiltonian on a fixed sparsity pattern,
H(theta) = sum_k theta_k B_k, so dL/dtheta = B^T dL/dH. The spectral window
is held fixed at the value given (it must contain the spectrum for every
theta the caller evaluates), and for an overlap the factor Z is held fixed.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import _lib
from .bsr import BSR
from .lattice import Edges, Structure


@dataclass
class LinearOperatorModel:
    """H(theta) = sum_k theta_k B_k; basis: [nnzb * b * b, n_params] dense,
    rows in the pattern's entry order (row-major within blocks)."""

    pattern: BSR
    basis: np.ndarray
    names: tuple

    def build(self, theta) -> BSR:
        vals = self.basis @ np.asarray(theta, dtype=np.float64)
        b = self.pattern.block_dim
        return BSR(self.pattern.n_block_rows, b, self.pattern.row_ptr, self.pattern.col_ind,
                   vals.reshape(-1, b, b))

    def contract(self, grad_entries: np.ndarray) -> np.ndarray:
        return self.basis.T @ grad_entries.reshape(-1)


def sk_pz_parameterization(st: Structure, edges: Edges, shell_radii) -> LinearOperatorModel:
    """p_z model with parameters (e0, t_1 .. t_K): onsite e0 on every atom,
    pp_pi hopping t_k on bonds of shell k (d <= shell_radii[k], first match),
    Slater-Koster factor 1 - n^2 (n = z / d). Zero-coefficient entries stay in
    the pattern so every theta shares it."""
    radii = list(shell_radii)
    n = st.n_atoms
    shell = np.full(edges.src.size, -1)
    for k, r in sorted(enumerate(radii), key=lambda t: t[1], reverse=True):
        shell[edges.dist <= r] = k
    keep = shell >= 0
    src, dst, sh = edges.src[keep], edges.dst[keep], shell[keep]
    n2 = (edges.vec[keep, 2] / edges.dist[keep]) ** 2
    rows = np.concatenate([np.arange(n), src])
    cols = np.concatenate([np.arange(n), dst])
    P = 1 + len(radii)
    cols_basis = []
    pattern = None
    for p in range(P):
        coef = np.zeros(rows.size)
        if p == 0:
            coef[:n] = 1.0
        else:
            coef[n:] = np.where(sh == p - 1, 1.0 - n2, 0.0)
        Hp = BSR.from_blocks(n, 1, rows, cols, coef.reshape(-1, 1, 1))
        pattern = Hp if pattern is None else pattern
        if not (np.array_equal(Hp.row_ptr, pattern.row_ptr) and np.array_equal(Hp.col_ind, pattern.col_ind)):
            raise RuntimeError("parameter patterns differ")
        cols_basis.append(Hp.values.reshape(-1))
    names = ("e0",) + tuple(f"t{k + 1}" for k in range(len(radii)))
    return LinearOperatorModel(pattern, np.stack(cols_basis, axis=1), names)


class DifferentiableKPM:
    """Moments mu(theta) on one device context with a backward pass."""

    def __init__(self, model: LinearOperatorModel, nsamples: int, window, n_moments: int,
                 checkpoint_stride: int = 32, sample_kind: str = "rademacher", seed: int = 20240917,
                 Z: BSR | None = None, device: int = 0):
        self.model = model
        self.M = int(n_moments)
        emin, emax = window
        if not (math.isfinite(emin) and math.isfinite(emax) and emin < emax):
            raise RuntimeError("window must be finite with emin < emax")
        self.window = (float(emin), float(emax))
        self.ctx = _lib.Context(device)
        rows = model.pattern.n_rows
        kind = _lib.SAMPLES_PHASE if sample_kind == "phase" else _lib.SAMPLES_RADEMACHER
        self.ctx.set_samples(rows, nsamples, kind, seed)
        form = _lib.FORM_ORTHOGONAL
        if Z is not None:
            self.ctx.upload_bsr(_lib.SLOT_Z, Z)
            self.ctx.upload_bsr(_lib.SLOT_ZH, Z.conj_transpose())
            form = _lib.FORM_SANDWICH
        self.form = form
        self.ctx.set_checkpointing(checkpoint_stride)
        self._theta = None

    def forward(self, theta) -> np.ndarray:
        theta = np.asarray(theta, dtype=np.float64)
        self.ctx.upload_bsr(_lib.SLOT_H, self.model.build(theta))
        emin, emax = self.window
        self.ctx.set_operator(self.form, (emax + emin) / 2.0, (emax - emin) / 2.0)
        mu = self.ctx.kpm_moments(self.M, _lib.PREC_FP64, _lib.DOUBLING_ON)
        self._theta = theta.copy()
        return mu.real.copy()

    def backward(self, dL_dmu) -> np.ndarray:
        """dL/dtheta for the last forward()."""
        if self._theta is None:
            raise RuntimeError("backward() needs a forward() first")
        g = self.ctx.kpm_moments_backward(np.asarray(dL_dmu, dtype=np.float64))
        return self.model.contract(g)

    def moments(self, theta):
        """torch entry point: mu(theta) with autograd."""
        return _kpm_fn().apply(theta, self)


_FN = None


def _kpm_fn():
    global _FN
    if _FN is not None:
        return _FN
    import torch

    class KPMMoments(torch.autograd.Function):
        @staticmethod
        def forward(ctx, theta, engine):
            mu = engine.forward(theta.detach().cpu().double().numpy())
            ctx.engine = engine
            return torch.from_numpy(mu)

        @staticmethod
        def backward(ctx, grad_mu):
            g = ctx.engine.backward(grad_mu.detach().cpu().double().numpy())
            return torch.from_numpy(g), None

    _FN = KPMMoments
    return _FN


def kpm_dos_torch(mu, x, edge_inset: float):
    """2 rho(x) with Jackson weights (rsatb/calc/kpm.hpp:206-241 and the
    ComputeDos 2 rho convention, observables.hpp:388-393), differentiable in mu."""
    import torch

    M = mu.shape[0]
    xs = np.clip(np.asarray(x, dtype=np.float64), -1.0 + edge_inset, 1.0 - edge_inset)
    denom = M + 1.0
    n = np.arange(M)
    gk = ((denom - n) * np.cos(np.pi * n / denom) + np.sin(np.pi * n / denom) / np.tan(np.pi / denom)) / denom
    gk[0] = 1.0
    w = 2.0 * gk
    w[0] = gk[0]
    T = np.cos(np.outer(n, np.arccos(xs)))  # T_n(x) = cos(n acos x) on [-1, 1]
    A = (w[:, None] * T) / (np.pi * np.sqrt(np.maximum(1e-14, 1.0 - xs * xs)))[None, :]
    return 2.0 * (mu @ torch.from_numpy(A))
