"""Differentiable KPM moments: a torch.autograd.Function over the library's
adjoint Chebyshev recurrence (rg_kpm_moments_backward_ex); and a
differentiable DC conductivity spectrum over its tangent recurrence
(rg_dc_spectrum_tangent).

PyTorch is used only here, as the autograd wrapper: the forward and the
backward both run in the gfx950 library; torch sees the moments (a real
vector) and hands back dL/dmu. Parameters enter through affine
parameterizations on fixed sparsity patterns, H(theta) = sum_k theta_k B_k
and, for a non-orthogonal basis, S(theta) = S0 + sum_k theta_k C_k, so
dL/dtheta = B^T dL/dH + C^T dL/dS; dL/dS runs through the graph-local factor
Z = S^{-1/2} the library builds on the device (rg_isqrt_patches). The spectral
window is held fixed at the value given (it must contain the spectrum for
every theta the caller evaluates).
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
    """A(theta) = offset + sum_k theta_k B_k; basis: [nnzb * b * b, n_params]
    dense, rows in the pattern's entry order (row-major within blocks);
    offset (optional, same layout) the constant part, e.g. an overlap's
    identity."""

    pattern: BSR
    basis: np.ndarray
    names: tuple
    offset: np.ndarray | None = None

    def build(self, theta) -> BSR:
        vals = self.basis @ np.asarray(theta, dtype=np.float64)
        if self.offset is not None:
            vals = vals + self.offset
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


def sk_pz_overlap_parameterization(st: Structure, edges: Edges, shell_radii) -> LinearOperatorModel:
    """p_z overlap S = I + sum_k s_k (1 - n^2) on the bonds of shell k
    (d <= shell_radii[k], first match; n = z / d), the Slater-Koster pp_pi
    overlap with unit onsite overlap: the affine model of the learnable
    nearest-neighbour overlap s1."""
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
    coef0 = np.zeros(rows.size)
    coef0[:n] = 1.0
    S0 = BSR.from_blocks(n, 1, rows, cols, coef0.reshape(-1, 1, 1))
    cols_basis = []
    for k in range(len(radii)):
        coef = np.zeros(rows.size)
        coef[n:] = np.where(sh == k, 1.0 - n2, 0.0)
        Sk = BSR.from_blocks(n, 1, rows, cols, coef.reshape(-1, 1, 1))
        if not (np.array_equal(Sk.row_ptr, S0.row_ptr) and np.array_equal(Sk.col_ind, S0.col_ind)):
            raise RuntimeError("parameter patterns differ")
        cols_basis.append(Sk.values.reshape(-1))
    names = tuple(f"s{k + 1}" for k in range(len(radii)))
    return LinearOperatorModel(S0, np.stack(cols_basis, axis=1), names, offset=S0.values.reshape(-1).copy())


@dataclass
class LinearTransportModel:
    """H(theta) = sum_k theta_k H_k with the bond-moment velocity factors
    B_a(theta) = sum_k theta_k B_{a,k} (V_a = i B_a; B_{a,k} holds the bond
    vector's a component times H_k's edge entries, onsite entries 0), on one
    pattern. Built by linear_transport_model."""

    hamiltonian: LinearOperatorModel
    velocity: tuple  # three LinearOperatorModel, one per axis


def linear_transport_model(n_atoms: int, edges: Edges, edge_coef, onsite_coef, names) -> LinearTransportModel:
    """One parameter k per column: edge_coef [E, P] is theta_k's hopping
    coefficient on each directed edge (src -> dst, the p_z entry H_ij), and
    onsite_coef [n_atoms, P] its onsite coefficient. Entries of periodic
    images that land on one (i, j) are summed, for H and for every B_a with
    its own bond vector, as the velocity of an orthogonal basis at Gamma
    requires (rsatb/operator/commutator_velocity.hpp:7-43)."""
    ec = np.asarray(edge_coef, dtype=np.float64).reshape(edges.src.size, -1)
    oc = np.asarray(onsite_coef, dtype=np.float64).reshape(n_atoms, -1)
    if ec.shape[1] != oc.shape[1] or ec.shape[1] != len(names):
        raise RuntimeError("edge and onsite coefficients need one column per parameter name")
    rows = np.concatenate([np.arange(n_atoms), edges.src])
    cols = np.concatenate([np.arange(n_atoms), edges.dst])
    weights = [None, edges.vec[:, 0], edges.vec[:, 1], edges.vec[:, 2]]
    models = []
    pattern = None
    for w in weights:
        basis = []
        for k in range(ec.shape[1]):
            if w is None:
                coef = np.concatenate([oc[:, k], ec[:, k]])
            else:
                coef = np.concatenate([np.zeros(n_atoms), ec[:, k] * w])
            Hk = BSR.from_blocks(n_atoms, 1, rows, cols, coef.reshape(-1, 1, 1))
            pattern = Hk if pattern is None else pattern
            if not (np.array_equal(Hk.row_ptr, pattern.row_ptr) and np.array_equal(Hk.col_ind, pattern.col_ind)):
                raise RuntimeError("parameter patterns differ")
            basis.append(Hk.values.reshape(-1))
        models.append(LinearOperatorModel(pattern, np.stack(basis, axis=1), tuple(names)))
    return LinearTransportModel(models[0], tuple(models[1:]))


def tblg_transport_model(reps: int):
    """Twisted bilayer graphene at 1.05 degrees ((31, 32) commensurate cell,
    11,908 atoms) tiled reps x reps, one p_z orbital per atom, and the two-
    parameter hopping model H = theta_intra H_intra + theta_inter H_inter:
    t(r) = Vpp_pi(r) (1 - n^2) + Vpp_sigma(r) n^2 with Vpp_pi = -2.7
    exp(-(r - 1.42) / 0.319) and Vpp_sigma = 0.48 exp(-(r - 3.35) / 0.319) eV,
    n the bond's z cosine, on every pair within 3.8 Angstrom; orthogonal
    basis."""
    from .lattice import neighbor_edges, tile, twisted_bilayer_graphene

    st = tile(twisted_bilayer_graphene(31, 32), (reps, reps, 1))
    e = neighbor_edges(st, 3.8)
    inter = st.atom_type[e.src] != st.atom_type[e.dst]
    n2 = (e.vec[:, 2] / e.dist) ** 2
    t_r = -2.7 * np.exp(-(e.dist - 1.42) / 0.319) * (1 - n2) + 0.48 * np.exp(-(e.dist - 3.35) / 0.319) * n2
    ec = np.zeros((e.src.size, 2))
    ec[~inter, 0] = t_r[~inter]
    ec[inter, 1] = t_r[inter]
    model = linear_transport_model(st.n_atoms, e, ec, np.zeros((st.n_atoms, 2)), ("intra", "inter"))
    return st, e, model



class DifferentiableConductivity:
    """The DC spectrum sigma(theta) (dc_spectrum_fftcheb,
    rsatb/calc/fftcheb_transport.hpp:543-653, orthogonal form at Gamma) and
    its Jacobian d sigma / d theta, one tangent pass per parameter
    (rg_dc_spectrum_tangent): forward-mode, which for the few parameters of a
    hopping parameterization costs (1 + P) spectra instead of an adjoint's
    stored cotangent sequences. The window is held fixed."""

    def __init__(self, model: LinearTransportModel, nsamples: int, window, n_moments: int,
                 formulation: str = "greenwood", kernel: str = "jackson", prefactor: float = 1.0,
                 sample_kind: str = "rademacher", seed: int = 20240917, device: int = 0):
        forms = {"greenwood": 0, "bastin": 1}
        kernels = {"jackson": _lib.KERNEL_JACKSON, "costaper": _lib.KERNEL_COSTAPER, "none": _lib.KERNEL_NONE}
        if formulation not in forms:
            raise RuntimeError(f"Unknown dc formulation: {formulation}")
        if kernel not in kernels:
            raise RuntimeError(f"unknown kpm kernel '{kernel}'")
        emin, emax = window
        if not (math.isfinite(emin) and math.isfinite(emax) and emin < emax):
            raise RuntimeError("window must be finite with emin < emax")
        self.model = model
        self.M = int(n_moments)
        self.formulation = forms[formulation]
        self.kernel = kernels[kernel]
        self.prefactor = float(prefactor)
        self.ctx = _lib.Context(device)
        kind = _lib.SAMPLES_PHASE if sample_kind == "phase" else _lib.SAMPLES_RADEMACHER
        self.ctx.set_samples(model.hamiltonian.pattern.n_rows, nsamples, kind, seed)
        self.ctx.set_operator(_lib.FORM_ORTHOGONAL, (emax + emin) / 2.0, (emax - emin) / 2.0)
        self.energy = None

    def _bind(self, theta):
        self.ctx.upload_bsr(_lib.SLOT_H, self.model.hamiltonian.build(theta))
        for a in range(3):
            self.ctx.upload_bsr(_lib.SLOT_VX + a, self.model.velocity[a].build(theta))

    def spectrum(self, theta) -> np.ndarray:
        """sigma [M, 6] on the ascending nodes (energies in self.energy)."""
        theta = np.asarray(theta, dtype=np.float64)
        self._bind(theta)
        sigma, self.energy = self.ctx.dc_spectrum(self.M, self.formulation, self.kernel, self.prefactor)
        return sigma

    def jacobian(self, theta):
        """(sigma [M, 6], d sigma / d theta [P, M, 6])."""
        theta = np.asarray(theta, dtype=np.float64)
        self._bind(theta)
        P = theta.size
        jac = np.zeros((P, self.M, 6))
        sigma = None
        for k in range(P):
            e_k = np.zeros(P)
            e_k[k] = 1.0
            self.ctx.upload_bsr(_lib.SLOT_DH, self.model.hamiltonian.build(e_k))
            for a in range(3):
                self.ctx.upload_bsr(_lib.SLOT_DVX + a, self.model.velocity[a].build(e_k))
            sigma, jac[k], self.energy = self.ctx.dc_spectrum_tangent(self.M, self.formulation, self.kernel,
                                                                      self.prefactor)
        if sigma is None:
            sigma = self.spectrum(theta)
        return sigma, jac

    def conductivity(self, theta):
        """torch entry point: sigma(theta) [M, 6] with autograd."""
        return _cond_fn().apply(theta, self)


class DifferentiableKPM:
    """Moments mu(theta) on one device context with a backward pass.

    theta = (H parameters of `model`, then the overlap parameters of
    `overlap` when given). With an overlap model every forward builds S(theta)
    and its graph-local factor Z on the device (rg_isqrt_patches) and runs the
    sandwich form; the backward returns dL/dtheta through H and through S -> Z
    (rg_kpm_moments_backward_ex). A fixed factor Z (`Z`) gives the sandwich
    form with no overlap gradient. Every forward is stamped: a backward for an
    earlier forward (two nodes of one torch graph) re-runs that forward first,
    so each node's gradient uses its own states."""

    def __init__(self, model: LinearOperatorModel, nsamples: int, window, n_moments: int,
                 checkpoint_stride: int = 32, sample_kind: str = "rademacher", seed: int = 20240917,
                 Z: BSR | None = None, overlap: LinearOperatorModel | None = None, isqrt_tol: float = 1e-12,
                 isqrt_maxit: int = 60, device: int = 0):
        if Z is not None and overlap is not None:
            raise RuntimeError("give a fixed factor Z or an overlap model, not both")
        self.model = model
        self.overlap = overlap
        self.M = int(n_moments)
        emin, emax = window
        if not (math.isfinite(emin) and math.isfinite(emax) and emin < emax):
            raise RuntimeError("window must be finite with emin < emax")
        self.window = (float(emin), float(emax))
        self.isqrt = (float(isqrt_tol), int(isqrt_maxit))
        self.ctx = _lib.Context(device)
        rows = model.pattern.n_rows
        kind = _lib.SAMPLES_PHASE if sample_kind == "phase" else _lib.SAMPLES_RADEMACHER
        self.ctx.set_samples(rows, nsamples, kind, seed)
        form = _lib.FORM_ORTHOGONAL
        if Z is not None:
            self.ctx.upload_bsr(_lib.SLOT_Z, Z)
            self.ctx.upload_bsr(_lib.SLOT_ZH, Z.conj_transpose())
            form = _lib.FORM_SANDWICH
        if overlap is not None:
            if overlap.pattern.n_rows != rows:
                raise RuntimeError("the overlap model and the Hamiltonian model differ in rows")
            form = _lib.FORM_SANDWICH
        self.form = form
        self.ctx.set_checkpointing(checkpoint_stride)
        self._theta = None
        self._stamp = 0

    @property
    def n_params(self) -> int:
        return self.model.basis.shape[1] + (0 if self.overlap is None else self.overlap.basis.shape[1])

    def _split(self, theta):
        nh = self.model.basis.shape[1]
        return theta[:nh], theta[nh:]

    def forward(self, theta) -> np.ndarray:
        theta = np.asarray(theta, dtype=np.float64)
        if theta.size != self.n_params:
            raise RuntimeError(f"theta has {theta.size} entries; the models take {self.n_params}")
        th_h, th_s = self._split(theta)
        if self.overlap is not None:
            self.ctx.upload_bsr(_lib.SLOT_S, self.overlap.build(th_s))
            self.ctx.isqrt_patches(*self.isqrt)
        self.ctx.upload_bsr(_lib.SLOT_H, self.model.build(th_h))
        emin, emax = self.window
        self.ctx.set_operator(self.form, (emax + emin) / 2.0, (emax - emin) / 2.0)
        mu = self.ctx.kpm_moments(self.M, _lib.PREC_FP64, _lib.DOUBLING_ON)
        self._theta = theta.copy()
        self._stamp += 1
        return mu.real.copy()

    def backward(self, dL_dmu, stamp: int | None = None, theta=None) -> np.ndarray:
        """dL/dtheta for the last forward(), or for the forward `stamp` (at
        `theta`), re-run first when a later forward replaced its states."""
        if self._theta is None:
            raise RuntimeError("backward() needs a forward() first")
        if stamp is not None and stamp != self._stamp:
            if theta is None:
                raise RuntimeError("a backward for an earlier forward needs that forward's theta")
            self.forward(theta)
        g = np.asarray(dL_dmu, dtype=np.float64)
        if self.overlap is None:
            return self.model.contract(self.ctx.kpm_moments_backward(g))
        outs = self.ctx.kpm_moments_backward_ex(g, want_h=True, want_s=True)
        return np.concatenate([self.model.contract(outs["H"]), self.overlap.contract(outs["S"])])

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
            th = theta.detach().cpu().double().numpy()
            mu = engine.forward(th)
            ctx.engine = engine
            ctx.stamp = engine._stamp
            ctx.theta = th.copy()
            return torch.from_numpy(mu)

        @staticmethod
        def backward(ctx, grad_mu):
            g = ctx.engine.backward(grad_mu.detach().cpu().double().numpy(), stamp=ctx.stamp, theta=ctx.theta)
            return torch.from_numpy(g), None

    _FN = KPMMoments
    return _FN


_COND_FN = None


def _cond_fn():
    global _COND_FN
    if _COND_FN is not None:
        return _COND_FN
    import torch

    class DcSpectrum(torch.autograd.Function):
        @staticmethod
        def forward(ctx, theta, engine):
            sigma, jac = engine.jacobian(theta.detach().cpu().double().numpy())
            ctx.jac = jac
            return torch.from_numpy(sigma)

        @staticmethod
        def backward(ctx, grad_sigma):
            g = np.tensordot(ctx.jac, grad_sigma.detach().cpu().double().numpy(), axes=([1, 2], [0, 1]))
            return torch.from_numpy(g), None

    _COND_FN = DcSpectrum
    return _COND_FN


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
