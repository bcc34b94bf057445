"""Python API mirroring rsatb's driver (rsatb/driver.py) for the KPM DOS path.

    calc = RSATB(structure, model, SetupOptions(nsamples=16, energy_bounds=(-8.6, 8.6)))
    dos, energy = calc.compute_dos(dE=0.05)

or, from a pre-built neighbour graph (rsatb/driver.py:73-115,
rsatb/pybind_rsatb.cpp:566-593):

    calc = RSATB.from_graph(atom_index, atom_type, edge_index, edge_shift_vec,
                            cell, pos, model, options)

Names, argument meanings, defaults and error behaviour follow the reference:
SetupOptions / PreconditionerOptions carry the same fields and defaults as
their pybind bindings (rsatb/pybind_rsatb.cpp:229-341); compute_dos returns
(dos, energy) with dos holding all k-points in one flat array
(rsatb/driver.py:120-136); configuration errors raise RuntimeError eagerly
(rsatb/calc/observables.hpp:316-324); an interrupted run raises
KeyboardInterrupt (rsatb/pybind_rsatb.cpp:92-93). All arithmetic runs in the
gfx950 library; this module only assembles operands and calls it.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib, transport
from .bsr import BSR
from .lattice import Structure, edges_from_graph, neighbor_edges
from .models import assemble
from .pdos import pdos_groups_by_atom, pdos_groups_by_orbital_slot


@dataclass
class PreconditionerOptions:
    """rsatb PreconditionerOptions (rsatb/pybind_rsatb.cpp:229-261)."""

    operator_form: str = "solve"
    tol: float = 1e-7
    sptol: float = 1e-5
    maxiter: int = 0

    def __post_init__(self):
        if self.operator_form not in ("solve", "sandwich"):
            raise RuntimeError(f"Unknown operator form '{self.operator_form}' (expected solve or sandwich)")


@dataclass
class SetupOptions:
    """rsatb SetupOptions (rsatb/pybind_rsatb.cpp:263-341), plus the sample
    distribution: "phase" (unit-norm random phase, the reference's draw,
    rsatb/calc/observables.hpp:54-97) or "rademacher"."""

    kpoints: list = field(default_factory=lambda: [[0.0, 0.0, 0.0]])
    nsamples: int = 1
    sample_seed: int = 20240917
    compute_bounds: bool = True
    energy_bounds: tuple | None = None
    bounds_zero_tol: float = 1e-12
    bounds_max_matvecs: int = 10000
    preconditioner: PreconditionerOptions = field(default_factory=PreconditionerOptions)
    sample_kind: str = "phase"
    precision: str = "fp64"

    def __post_init__(self):
        if self.energy_bounds is not None:
            b = list(self.energy_bounds)
            if len(b) != 2:
                raise RuntimeError("energy_bounds must contain exactly two values: [emin, emax].")
            self.energy_bounds = (float(b[0]), float(b[1]))
        if self.sample_kind not in ("phase", "rademacher"):
            raise RuntimeError("sample_kind must be 'phase' or 'rademacher'")
        if self.precision not in ("fp64", "fp32"):
            raise RuntimeError("precision must be 'fp64' or 'fp32'")


def make_explicit_bounds(emin: float, emax: float) -> tuple[float, float]:
    """MakeExplicitBounds (rsatb/setup/spectral_bounds.hpp:427-432)."""
    if not (math.isfinite(emin) and math.isfinite(emax) and emin < emax):
        raise RuntimeError("Explicit energy bounds must be finite and satisfy emin < emax.")
    return float(emin), float(emax)


class RSATB:
    """A calculation on one GPU: operators per k-point, samples, the DOS task."""

    def __init__(self, structure: Structure, model, options: SetupOptions | None = None, npar: int = 1,
                 device: int = 0, edges=None):
        if options is None:
            options = SetupOptions()
        if npar != 1:
            raise RuntimeError("npar > 1 (k-point groups across processes) is not provided on one GPU")
        if options.nsamples < 1:
            raise RuntimeError("nsamples must be positive")
        self.structure = structure
        self.model = model
        self.options = options
        self.edges = edges if edges is not None else neighbor_edges(structure, model.cutoff)
        self._ctx = _lib.Context(device)
        self._kstates = []
        self._bounds = None
        self.atom_index = np.arange(structure.n_atoms)
        self._setup()

    @classmethod
    def from_graph(cls, atom_index, atom_type, edge_index, edge_shift_vec, cell, pos, model,
                   options: SetupOptions | None = None, N_atom: int = 0, atom_offset: int = 0,
                   N_edge: int = 0, npar: int = 1, device: int = 0) -> "RSATB":
        """rsatb RSATB.from_graph: one process owns the whole graph here, so the
        local counts must equal the totals (rsatb/rsatb.hpp:467-491)."""
        atom_index = np.asarray(atom_index, dtype=np.int64).reshape(-1)
        n_atom = atom_index.size
        n_edge = np.asarray(edge_index).size // 2
        if N_atom and N_atom != n_atom:
            raise RuntimeError(f"N_atom={N_atom} but {n_atom} atoms were passed (one process owns all atoms)")
        if N_edge and N_edge != n_edge:
            raise RuntimeError(f"N_edge={N_edge} but {n_edge} edges were passed")
        if atom_offset != 0:
            raise RuntimeError("atom_offset must be 0 on a single process")
        if not np.array_equal(np.sort(atom_index), np.arange(n_atom)):
            raise RuntimeError("atom_index must be a permutation of 0..N_atom-1")
        pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
        st = Structure(pos, np.asarray(cell, dtype=np.float64).reshape(3, 3), (True, True, True),
                       np.asarray(atom_type, dtype=np.int32))
        edges = edges_from_graph(st, edge_index, edge_shift_vec)
        calc = cls(st, model, options, npar, device, edges=edges)
        calc.atom_index = atom_index
        return calc

    # ------------------------------------------------------------- setup
    def _setup(self) -> None:
        o = self.options
        st, model = self.structure, self.model
        sandwich = model.has_overlap()
        if sandwich and o.preconditioner.operator_form != "sandwich":
            raise RuntimeError(
                "This model has an overlap: the GPU path iterates the sandwich form Z^H H Z only; "
                "set PreconditionerOptions(operator_form='sandwich')."
            )
        if sandwich and o.preconditioner.maxiter != 0:
            raise RuntimeError(
                "PreconditionerOptions.maxiter > 0 asks for the sparse Newton-Schulz refinement, which the GPU "
                "path does not provide; the factor is the graph-local seed (maxiter = 0)."
            )
        device_sk = getattr(model, "device_assembly", False)
        for k in o.kpoints:
            if device_sk:
                # built on the device when the k-point is bound (rg_sk_assemble)
                self._kstates.append({"k": tuple(k), "H": None, "S": None, "Z": None, "sk": True,
                                      "has_S": sandwich})
                continue
            H = assemble(st, self.edges, model, "H", k)
            S = assemble(st, self.edges, model, "S", k) if sandwich else None
            self._kstates.append({"k": tuple(k), "H": H, "S": S, "Z": None, "has_S": sandwich})
        self._bound_k = None
        if o.energy_bounds is not None:
            self._bounds = make_explicit_bounds(*o.energy_bounds)
        elif o.compute_bounds:
            lo, hi = np.inf, -np.inf
            for k_idx in range(len(self._kstates)):
                self._bind_factors(k_idx)
                e0, e1 = self._ctx.estimate_bounds(self._form(k_idx), steps=self.lanczos_steps,
                                                   seed=o.sample_seed, padding=0.0)
                lo, hi = min(lo, e0), max(hi, e1)
            # padding of EstimateSpectralBounds (rsatb/setup/spectral_bounds.hpp:75, 488)
            self._bounds = make_explicit_bounds(lo - self.bounds_padding, hi + self.bounds_padding)
        n_rows = st.n_atoms * model.norb
        kind = _lib.SAMPLES_PHASE if o.sample_kind == "phase" else _lib.SAMPLES_RADEMACHER
        self._ctx.set_samples(n_rows, o.nsamples, kind, o.sample_seed)

    lanczos_steps = 200
    bounds_padding = 0.5

    def _form(self, idx: int) -> int:
        return _lib.FORM_SANDWICH if self._kstates[idx]["has_S"] else _lib.FORM_ORTHOGONAL

    def _sk_assemble(self, slot: int, which: str, k) -> None:
        st, e = self.structure, self.edges
        self._ctx.sk_assemble(slot, self.model.arrays(which), st.atom_type, st.positions, st.cell,
                              np.stack([e.src, e.dst], axis=1), e.shift, None if not np.any(k) else k)

    def _bind_factors(self, idx: int) -> None:
        """Puts H(k) (and Z, Z^H for an overlap) on the device."""
        if self._bound_k == idx:
            return
        ks = self._kstates[idx]
        if ks.get("sk"):
            self._sk_assemble(_lib.SLOT_H, "H", ks["k"])
        else:
            self._ctx.upload_bsr(_lib.SLOT_H, ks["H"])
        if ks["has_S"]:
            if ks["Z"] is not None:
                self._ctx.upload_bsr(_lib.SLOT_Z, ks["Z"])
                self._ctx.upload_bsr(_lib.SLOT_ZH, ks["Z"].conj_transpose())
            else:
                if ks.get("sk"):
                    self._sk_assemble(_lib.SLOT_S, "S", ks["k"])
                else:
                    self._ctx.upload_bsr(_lib.SLOT_S, ks["S"])
                ks["isqrt"] = self._ctx.isqrt_patches()
        self._bound_k = idx

    def _bind_kpoint(self, idx: int) -> None:
        self._bind_factors(idx)
        emin, emax = self.energy_bounds()
        self._ctx.set_operator(self._form(idx), (emax + emin) / 2.0, (emax - emin) / 2.0)

    def set_overlap_factor(self, Z: BSR, k_idx: int = 0) -> None:
        """Installs a factor Z (applied as Z x, with Z^H as its adjoint) for a
        k-point, as ISQRTMPreconditioner's borrowing constructor does
        (rsatb/calc/preconditioner.hpp:194-196); without one the factor is
        built on the device from S (rg_isqrt_patches)."""
        self._kstates[k_idx]["Z"] = Z
        if self._bound_k == k_idx:
            self._bound_k = None

    def overlap_factor(self, k_idx: int = 0):
        """The factor Z of a k-point as the device holds it (downloaded)."""
        self._bind_factors(k_idx)
        if not self._kstates[k_idx]["has_S"]:
            raise RuntimeError("this model has no overlap")
        return self._ctx.download_bsr(_lib.SLOT_Z)

    # ------------------------------------------------------------- tasks
    def energy_bounds(self) -> tuple[float, float]:
        if self._bounds is None:
            raise RuntimeError(
                "Spectral bound is not computed yet. Set it true during Setup stage for random state method support."
            )
        return self._bounds

    def kpm_moments(self, n_moments: int, k_idx: int = 0, precision: str | None = None,
                    doubling: bool | None = None, per_column: bool = False):
        """kpm_dos_moments (rsatb/calc/dos.hpp:124-189) on one k-point."""
        self._bind_kpoint(k_idx)
        prec = _lib.PREC_FP32 if (precision or self.options.precision) == "fp32" else _lib.PREC_FP64
        dbl = _lib.DOUBLING_AUTO if doubling is None else (_lib.DOUBLING_ON if doubling else _lib.DOUBLING_OFF)
        return self._ctx.kpm_moments(n_moments, prec, dbl, per_column)

    def compute_dos(self, dE: float, method: str = "kpm"):
        """ComputeDos (rsatb/calc/observables.hpp:311-444): returns (dos, energy),
        dos flat over k-points."""
        if method != "kpm":
            raise RuntimeError(f"dos method must be 'kpm' on the GPU path, got '{method}'")
        if not (dE > 0.0):
            raise RuntimeError("dos dE must be a positive energy resolution in eV")
        emin, emax = self.energy_bounds()
        prec = _lib.PREC_FP32 if self.options.precision == "fp32" else _lib.PREC_FP64
        parts, omega = [], None
        for k_idx in range(len(self._kstates)):
            self._bind_kpoint(k_idx)
            dos, omega = self._ctx.compute_dos(emin, emax, dE, prec)
            parts.append(dos)
        return np.concatenate(parts), omega

    def compute_pdos(self, dE: float, projection: str = "atom"):
        """ComputePdos (rsatb/calc/observables.hpp:446-560) and its Python
        return value (rsatb/pybind_rsatb.cpp:671-698): a dict with energy
        (2nt,), pdos (n_k, n_group, 2nt), total (n_k, 2nt), n_group,
        population, group_type and group_slot. Both operator forms here are
        Hermitian, so the population is always 'lowdin'."""
        if not (dE > 0.0):
            raise RuntimeError("pdos dE must be a positive energy resolution in eV")
        emin, emax = self.energy_bounds()
        st, norb = self.structure, self.model.norb
        type_norb = np.full(int(st.atom_type.max()) + 1 if st.n_atoms else 1, norb)
        if projection == "atom":
            groups = pdos_groups_by_atom(st.atom_type, type_norb, self.atom_index)
        elif projection == "orbital_slot":
            groups = pdos_groups_by_orbital_slot(st.atom_type, type_norb)
        else:
            raise RuntimeError(f"pdos projection must be 'atom' or 'orbital_slot', got '{projection}'")
        nt, N = _lib.dos_plan(emin, emax, dE)
        nt //= 2
        x = np.concatenate([(np.arange(nt) - nt) / nt, np.arange(nt) / nt])
        half, shift = (emax - emin) / 2.0, (emax + emin) / 2.0
        prec = _lib.PREC_FP32 if self.options.precision == "fp32" else _lib.PREC_FP64
        nk = len(self._kstates)
        pdos = np.zeros((nk, groups.n_groups, 2 * nt))
        total = np.zeros((nk, 2 * nt))
        for k_idx in range(nk):
            self._bind_kpoint(k_idx)
            mu_t, mu_g = self._ctx.kpm_pdos_moments(N, groups.group_of_row, groups.n_groups, prec)
            total[k_idx] = 2.0 * _lib.kpm_density(mu_t, x, 0.5 / nt)
            for g in range(groups.n_groups):
                pdos[k_idx, g] = 2.0 * _lib.kpm_density(mu_g[g], x, 0.5 / nt)
        energy = np.concatenate([(np.arange(nt) - nt) * half / nt + shift, np.arange(nt) * half / nt + shift])
        return {
            "energy": energy, "pdos": pdos, "total": total, "n_group": groups.n_groups,
            "population": "lowdin", "group_type": groups.group_type, "group_slot": groups.group_slot,
        }

    def transport_volume(self, axis: str = "abc") -> float:
        """TransportNormalizationVolume (rsatb/calc/observables.hpp:277-285):
        a cell-vector length for one axis, the spanned area for two, the cell
        volume for "abc"."""
        cell = self.structure.cell
        rows = ["abc".index(ch) for ch in axis] if axis and all(ch in "abc" for ch in axis) else None
        if rows is None or len(set(rows)) != len(rows):
            raise RuntimeError(f"unknown normalization axis '{axis}'")
        if len(rows) == 1:
            return float(np.linalg.norm(cell[rows[0]]))
        if len(rows) == 2:
            return float(np.linalg.norm(np.cross(cell[rows[0]], cell[rows[1]])))
        return float(abs(np.linalg.det(cell)))

    def compute_dc_spectrum(self, n_moments: int, formulation: str = "greenwood", axis: str = "abc",
                            partition_sections: int = 0, max_workspace_mb: int = 0, kernel: str = "jackson",
                            method: str = "fftcheb"):
        """compute_dc_spectrum (rsatb/driver.py:366-396, observables.hpp:691-779):
        the zero-temperature Greenwood conductivity on the Chebyshev nodes,
        sigma (n_k, n_moments, 6) in S/m with both spins, components (xx, xy,
        xz, yy, yz, zz) (Greenwood fills the diagonal; Bastin all six, as the
        T = 0 conductivity against the Fermi energy), and the energies in eV. The left and right sequences are kept whole on
        the device (partition_sections and max_workspace_mb are accepted and
        do not apply). At a k-point the velocity carries the same Bloch phase
        as H(k)."""
        forms = {"greenwood": 0, "bastin": 1}
        if formulation not in forms:
            raise RuntimeError(f"Unknown dc formulation: {formulation}")
        if method not in ("fftcheb", "qe"):
            raise RuntimeError(f"dc method must be 'fftcheb' or 'qe', got '{method}'")
        if method == "qe" and formulation != "greenwood":
            raise RuntimeError("dc method 'qe' supports formulation 'greenwood' only; "
                               "'bastin' needs the fftcheb row transform")
        kernels = {"jackson": _lib.KERNEL_JACKSON, "costaper": _lib.KERNEL_COSTAPER, "none": _lib.KERNEL_NONE}
        if kernel not in kernels:
            raise RuntimeError(f"unknown kpm kernel '{kernel}'")
        if not getattr(self.model, "device_assembly", False) or self.model.has_overlap():
            raise RuntimeError("compute_dc_spectrum requires a velocity-capable model/operator setup: an "
                               "orthogonal SktbModel (the overlap's composite velocity is not provided)")
        if n_moments < 1:
            raise RuntimeError("n_moments must be positive")
        emin, emax = self.energy_bounds()
        st, e = self.structure, self.edges
        n_rows = st.n_atoms * self.model.norb
        # the precision form (SetupOptions.precision "fp32"): FP32 recurrence, FP64 moments
        prec = _lib.PREC_FP32 if self.options.precision == "fp32" else _lib.PREC_FP64
        # pi g_s (4 Greenwood, 8 Bastin) N / volume (observables.hpp:722-733)
        prefactor = np.pi * 2.0 * (4.0 if formulation == "greenwood" else 8.0) * n_rows / self.transport_volume(axis)
        sigma = np.zeros((len(self._kstates), n_moments, 6))
        energy = None
        for k_idx, ks in enumerate(self._kstates):
            self._bind_kpoint(k_idx)
            # the velocity in the gauge of the bound H(k) (Bloch phase per edge)
            self._ctx.sk_velocity(self.model.arrays("H"), st.atom_type, st.positions, st.cell,
                                  np.stack([e.src, e.dst], axis=1), e.shift,
                                  kpoint=ks["k"] if np.any(ks["k"]) else None)
            sigma[k_idx], energy = self._ctx.dc_spectrum(n_moments, forms[formulation], kernels[kernel], prefactor,
                                                         precision=prec)
        # e^2 / (hbar Angstrom) in S/m (rsatb/driver.py:393)
        return sigma * 2.4341348 * 1e6, energy

    def compute_dc(self, n_moments: int, mu: float, temp: float, formulation: str = "greenwood",
                   axis: str = "abc", partition_sections: int = 0, max_workspace_mb: int = 0,
                   kernel: str = "jackson", method: str = "fftcheb"):
        """compute_dc (rsatb/driver.py:398-429, observables.hpp:792-854): the
        conductivity tensor (n_k, 6) in S/m at one chemical potential mu (eV)
        and temperature temp (K), the spectrum of compute_dc_spectrum
        integrated against -df/dE per k-point. method "qe" keeps the spectrum
        only on the nodes within max(20 kT, 3 pi a / n_moments) of mu, zero
        elsewhere, as the reference's windowed route delivers it (the device
        evaluates every node either way; the nodes kept are identical)."""
        sigma, energy = self.compute_dc_spectrum(n_moments, formulation, axis, partition_sections,
                                                 max_workspace_mb, kernel, method)
        if method == "qe":
            emin, emax = self.energy_bounds()
            nodes = transport.fermi_window_nodes(n_moments, 0.5 * (emax + emin), 0.5 * (emax - emin), mu, temp)
            keep = np.zeros(n_moments, dtype=bool)
            keep[nodes] = True
            sigma = np.where(keep[None, :, None], sigma, 0.0)
        return np.stack([transport.integrate_dc_spectrum(sigma[k], energy, mu, temp)
                         for k in range(sigma.shape[0])])

    def dump_operator_coo(self, which: str, axis: int = 0, k_idx: int = 0):
        """Global COO triplets of H or S (rsatb/pybind_rsatb.cpp:612-670)."""
        ks = self._kstates[k_idx]
        M = ks.get(which)
        if which not in ("H", "S"):
            raise RuntimeError(f"unknown operator name: {which}")
        if ks.get("sk") and (which == "H" or ks["has_S"]):
            # assembled on the device: build it into its slot and read it back
            slot = _lib.SLOT_H if which == "H" else _lib.SLOT_S
            self._sk_assemble(slot, which, ks["k"])
            self._bound_k = None
            return self._ctx.download_bsr(slot).to_coo()
        if M is None:
            raise RuntimeError(f"{which} is not populated for this model")
        return M.to_coo()

    def cancel(self) -> None:
        """Asks a running moment loop to stop at its next step boundary."""
        self._ctx.cancel()

    @property
    def context(self) -> _lib.Context:
        return self._ctx
