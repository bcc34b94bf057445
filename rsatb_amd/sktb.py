"""Slater-Koster parameter tables for the device assembly of H and S.

The parameter model follows the reference's SK tight-binding input:
  - orbital vocabulary and channel names: rsatb/model/sktb/orbital_labels.hpp
    (shell aliases expand in DeePTB magnetic order, p -> py, pz, px and
    d -> dxy, dyz, dz2, dzx, dx2-y2, :45-60; channel names "pp_sigma" etc.
    with canonical shell order, :115-135);
  - radial functions: rsatb/model/sktb/radial.hpp (natural cubic spline
    through tabulated values, :93-131 and :139-172; exp_poly with a cosine
    taper, :174-195);
  - resolution into a table keyed by atom type: sk_table.hpp:276-424
    (species ordered by atomic number, onsite diagonals, each channel of a
    shell pair l < l' registered for (l, l') and (l', l), per-pair cutoffs).

`SkTable.arrays(which)` flattens one matrix ("H" or "S") of a resolved table
into the plain arrays the C-ABI entry `rg_sk_assemble` takes.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

_ELEMENTS = (
    "H He Li Be B C N O F Ne Na Mg Al Si P S Cl Ar K Ca Sc Ti V Cr Mn Fe Co Ni Cu Zn Ga Ge As Se Br Kr "
    "Rb Sr Y Zr Nb Mo Tc Ru Rh Pd Ag Cd In Sn Sb Te I Xe Cs Ba La Ce Pr Nd Pm Sm Eu Gd Tb Dy Ho Er Tm "
    "Yb Lu Hf Ta W Re Os Ir Pt Au Hg Tl Pb Bi Po At Rn"
).split()
_Z = {s: i + 1 for i, s in enumerate(_ELEMENTS)}

# (l, magnetic index) of every explicit orbital (orbital_labels.hpp:69-81)
ORBITALS = {
    "s": (0, 0),
    "py": (1, 0), "pz": (1, 1), "px": (1, 2),
    "dxy": (2, 0), "dyz": (2, 1), "dz2": (2, 2), "dzx": (2, 3), "dx2-y2": (2, 4),
}
_SHELLS = {"s": ["s"], "p": ["py", "pz", "px"], "d": ["dxy", "dyz", "dz2", "dzx", "dx2-y2"]}
_BONDS = {"sigma": 0, "pi": 1, "delta": 2}

RADIAL_SPLINE, RADIAL_EXP_POLY = 0, 1


def canonical_symbol(raw: str) -> str:
    s = raw.strip()
    s = s[:1].upper() + s[1:].lower()
    if s not in _Z:
        raise ValueError(f"unknown chemical species '{raw}'")
    return s


def canonical_orbital_label(raw: str) -> str:
    label = raw.strip().lower()
    return {"dxz": "dzx", "dx2y2": "dx2-y2"}.get(label, label)


def expand_basis_label(raw: str) -> list[str]:
    label = canonical_orbital_label(raw)
    if label in _SHELLS:
        return list(_SHELLS[label])
    if label in ORBITALS:
        return [label]
    raise ValueError(f"Unsupported SKTB orbital label: {raw}")


def parse_channel_name(raw: str) -> tuple[int, int, int]:
    """'pd_pi' -> (row_l, col_l, channel) = (1, 2, 1) (orbital_labels.hpp:115-135)."""
    name = raw.strip().lower()
    if len(name) < 4 or name[2] != "_" or name[0] not in "spd" or name[1] not in "spd":
        raise ValueError(f"Unsupported SKTB channel name: {raw}")
    row_l, col_l = "spd".index(name[0]), "spd".index(name[1])
    bond = name[3:]
    if bond not in _BONDS:
        raise ValueError(f"Unsupported SKTB channel kind: {bond}")
    ch = _BONDS[bond]
    if row_l > col_l:
        raise ValueError(f"SKTB channel name must use canonical shell order s, p, d: {raw}")
    if ch > min(row_l, col_l):
        raise ValueError(f"SKTB channel is not valid for shell pair: {raw}")
    return row_l, col_l, ch


def make_channel_name(shell_pair: str, bond: str) -> str:
    sp = shell_pair.strip().lower().replace("-", "").replace("_", "").replace(" ", "")
    name = f"{sp}_{bond.strip().lower()}"
    parse_channel_name(name)
    return name


@dataclass
class Radial:
    """Distance dependence of one channel (radial.hpp:22-196)."""

    kind: str
    r: np.ndarray = field(default_factory=lambda: np.zeros(0))
    values: np.ndarray = field(default_factory=lambda: np.zeros(0))
    coeffs: np.ndarray = field(default_factory=lambda: np.zeros(0))
    lambda_value: float = 0.0
    r_on: float = 0.0
    r_cut: float = 0.0
    y2: np.ndarray = field(default_factory=lambda: np.zeros(0))

    @staticmethod
    def spline(r, values, r_cut: float = 0.0) -> "Radial":
        r = np.asarray(r, dtype=np.float64)
        values = np.asarray(values, dtype=np.float64)
        if r.size == 0 or r.size != values.size:
            raise ValueError("A spline radial function requires equally sized non-empty r and values arrays.")
        rad = Radial("spline", r=r, values=values, r_cut=float(r_cut) if r_cut > 0 else float(r.max()))
        rad.validate("spline")
        rad.y2 = rad.spline_second_derivatives()
        return rad

    @staticmethod
    def exp_poly(coeffs, r_cut: float, lambda_value: float = 0.0, r_on: float = 0.0) -> "Radial":
        rad = Radial("exp_poly", coeffs=np.asarray(coeffs, dtype=np.float64), lambda_value=float(lambda_value),
                     r_on=float(r_on), r_cut=float(r_cut))
        rad.validate("exp_poly")
        return rad

    def validate(self, context: str) -> None:
        """radial.hpp:50-84."""
        if self.kind == "spline":
            if self.r.size == 0 or self.r.size != self.values.size:
                raise ValueError(f"{context}: spline radial function requires equally sized non-empty "
                                 "'r' and 'values' arrays.")
            d = np.diff(self.r)
            if np.any(np.abs(d) <= 1e-12) or np.any(d < 0):
                raise ValueError(f"{context}: spline radial distances must be strictly increasing.")
            if self.r_cut <= 0.0:
                raise ValueError(f"{context}: spline radial function requires a positive r_cut.")
        elif self.kind == "exp_poly":
            if self.coeffs.size == 0:
                raise ValueError(f"{context}: exp_poly radial function requires at least one coefficient.")
            if self.r_cut <= 0.0:
                raise ValueError(f"{context}: exp_poly radial function requires a positive r_cut.")
            if self.r_on < 0.0 or self.r_on > self.r_cut:
                raise ValueError(f"{context}: exp_poly radial function requires 0 <= r_on <= r_cut.")
        elif self.kind == "table":
            raise ValueError(f"{context}: radial function kind 'table' has been removed; use 'spline' "
                             "for tabulated data.")
        else:
            raise ValueError(f"{context}: unsupported radial function kind '{self.kind}'.")

    def spline_second_derivatives(self) -> np.ndarray:
        """Natural cubic spline second derivatives: the tridiagonal sweep of
        radial.hpp:93-125 (y2[0] = y2[n-1] = 0)."""
        r, v = self.r, self.values
        n = r.size
        if n < 2:
            raise ValueError("Spline radial function requires at least two tabulated points.")
        y2 = np.zeros(n)
        u = np.zeros(n - 1)
        for i in range(1, n - 1):
            hp, hn = r[i] - r[i - 1], r[i + 1] - r[i]
            sig = hp / (hp + hn)
            p = sig * y2[i - 1] + 2.0
            y2[i] = (sig - 1.0) / p
            u[i] = (6.0 * ((v[i + 1] - v[i]) / hn - (v[i] - v[i - 1]) / hp) / (hp + hn) - sig * u[i - 1]) / p
        for k in range(n - 2, -1, -1):
            y2[k] = y2[k] * y2[k + 1] + u[k]
        return y2


class SkParameters:
    """Onsite energies and pair channels of one matrix (sk_table.hpp:52-101)."""

    def __init__(self):
        self.onsite: dict[str, dict[str, float]] = {}
        self.pairs: list[tuple[str, str, dict[str, Radial]]] = []

    def enabled(self) -> bool:
        return bool(self.onsite) or bool(self.pairs)

    def set_onsite(self, species: str, orbital: str, value: float) -> "SkParameters":
        self.onsite.setdefault(canonical_symbol(species), {})[canonical_orbital_label(orbital)] = float(value)
        return self

    def pair(self, row_species: str, col_species: str) -> dict[str, Radial]:
        rs, cs = canonical_symbol(row_species), canonical_symbol(col_species)
        for r, c, ch in self.pairs:
            if r == rs and c == cs:
                return ch
        self.pairs.append((rs, cs, {}))
        return self.pairs[-1][2]

    def set_pair_channel(self, row_species: str, col_species: str, shell_pair: str, bond: str,
                         radial: Radial) -> "SkParameters":
        self.pair(row_species, col_species)[make_channel_name(shell_pair, bond)] = radial
        return self


@dataclass
class SkAtomType:
    symbol: str
    orbitals: list  # [(label, l, m)]
    h_onsite: np.ndarray
    s_onsite: np.ndarray | None = None


@dataclass
class SkPairChannel:
    row_l: int
    col_l: int
    channel: int
    h: Radial | None = None
    s: Radial | None = None


@dataclass
class SkAtomPair:
    channels: list = field(default_factory=list)
    h_max_cutoff: float = 0.0
    s_max_cutoff: float = 0.0


def _expand_onsite(raw: dict, orbitals: list, symbol: str, require_complete: bool, identity: bool) -> np.ndarray:
    """sk_table.hpp:212-250."""
    index = {lab: i for i, (lab, _, _) in enumerate(orbitals)}
    diag = np.full(len(orbitals), 1.0 if identity else 0.0)
    assigned = [identity] * len(orbitals)
    for key, value in raw.items():
        for lab in expand_basis_label(key):
            if lab not in index:
                raise ValueError(f"SKTB onsite specification for species {symbol} references orbital {lab} "
                                 "which is not in the declared basis.")
            if assigned[index[lab]] and not identity:
                raise ValueError(f"SKTB onsite specification for species {symbol} assigns orbital {lab} "
                                 "more than once.")
            diag[index[lab]] = value
            assigned[index[lab]] = True
    if require_complete:
        for i, ok in enumerate(assigned):
            if not ok:
                raise ValueError(f"SKTB onsite specification for species {symbol} is missing orbital "
                                 f"{orbitals[i][0]}.")
    return diag


class SkTable:
    """Resolved table keyed by atom type (sk_table.hpp:276-424)."""

    def __init__(self, basis: dict, hamiltonian: SkParameters, overlap: SkParameters | None = None):
        if not basis:
            raise ValueError("SKTB setup requires a basis definition.")
        if not hamiltonian.enabled():
            raise ValueError("SKTB setup requires Hamiltonian parameters.")
        declared = sorted(basis, key=lambda s: _Z[canonical_symbol(s)])
        symbols = [canonical_symbol(s) for s in declared]
        if len(set(symbols)) != len(symbols):
            raise ValueError("SKTB species must not contain duplicates.")
        self.has_overlap = overlap is not None and overlap.enabled()
        self.atom_types: list[SkAtomType] = []
        for raw, sym in zip(declared, symbols):
            orbitals, seen = [], set()
            for lab in basis[raw]:
                for e in expand_basis_label(lab):
                    if e in seen:
                        raise ValueError(f"SKTB basis for species {sym} contains a duplicate orbital: {e}")
                    seen.add(e)
                    orbitals.append((e,) + ORBITALS[e])
            if not orbitals:
                raise ValueError(f"SKTB basis for species {sym} must not be empty.")
            if sym not in hamiltonian.onsite:
                raise ValueError(f"SKTB Hamiltonian onsite terms are missing for species {sym}.")
            t = SkAtomType(sym, orbitals, _expand_onsite(hamiltonian.onsite[sym], orbitals, sym, True, False))
            if self.has_overlap:
                if sym in overlap.onsite:
                    t.s_onsite = _expand_onsite(overlap.onsite[sym], orbitals, sym, True, False)
                else:
                    t.s_onsite = np.ones(len(orbitals))
            self.atom_types.append(t)
        self.symbol_to_type = {t.symbol: i for i, t in enumerate(self.atom_types)}
        self.pairs: dict[tuple[int, int], SkAtomPair] = {}
        self._resolve(hamiltonian, False)
        if self.has_overlap:
            self._resolve(overlap, True)

    def _resolve(self, params: SkParameters, for_overlap: bool) -> None:
        name = "overlap" if for_overlap else "Hamiltonian"
        for rs, cs, channels in params.pairs:
            if rs not in self.symbol_to_type or cs not in self.symbol_to_type:
                raise ValueError(f"SKTB {name} pair ({rs}, {cs}) names a species that is not in the declared basis.")
            key = (self.symbol_to_type[rs], self.symbol_to_type[cs])
            pair = self.pairs.setdefault(key, SkAtomPair())
            for cname, radial in channels.items():
                row_l, col_l, ch = parse_channel_name(cname)
                radial.validate(f"SKTB {name} pair ({rs}, {cs}), channel {cname}")
                if radial.kind == "spline" and radial.y2.size == 0:
                    radial.y2 = radial.spline_second_derivatives()
                for a, b in ((row_l, col_l), (col_l, row_l)) if row_l != col_l else ((row_l, col_l),):
                    entry = next((c for c in pair.channels if (c.row_l, c.col_l, c.channel) == (a, b, ch)), None)
                    if entry is None:
                        entry = SkPairChannel(a, b, ch)
                        pair.channels.append(entry)
                    if (entry.s if for_overlap else entry.h) is not None:
                        raise ValueError(f"SKTB {name} pair ({rs}, {cs}) assigns channel '{cname}' more than once.")
                    if for_overlap:
                        entry.s = radial
                        pair.s_max_cutoff = max(pair.s_max_cutoff, radial.r_cut)
                    else:
                        entry.h = radial
                        pair.h_max_cutoff = max(pair.h_max_cutoff, radial.r_cut)

    @property
    def type_norb(self) -> list[int]:
        return [len(t.orbitals) for t in self.atom_types]

    def cutoff_radii(self) -> list[float]:
        """Half the largest bond cutoff touching each type (sk_table.hpp:176-191)."""
        radii = [0.0] * len(self.atom_types)
        for (rt, ct), p in self.pairs.items():
            r = 0.5 * max(p.h_max_cutoff, p.s_max_cutoff)
            radii[rt] = max(radii[rt], r)
            radii[ct] = max(radii[ct], r)
        return radii

    def arrays(self, which: str) -> dict:
        """Flat arrays of matrix `which` ("H" or "S") for rg_sk_assemble."""
        if which not in ("H", "S"):
            raise ValueError("which must be 'H' or 'S'")
        if which == "S" and not self.has_overlap:
            raise ValueError("SKTB overlap construction requires an overlap model.")
        nt = len(self.atom_types)
        norb = self.type_norb
        if len(set(norb)) != 1:
            raise ValueError("the device BSR holds uniform blocks: every atom type needs the same orbital count")
        b = norb[0]
        orb_l = np.array([[o[1] for o in t.orbitals] for t in self.atom_types], dtype=np.int32).ravel()
        orb_m = np.array([[o[2] for o in t.orbitals] for t in self.atom_types], dtype=np.int32).ravel()
        onsite = np.concatenate([t.h_onsite if which == "H" else t.s_onsite for t in self.atom_types])
        pair_cut = np.zeros(nt * nt)
        cols = {k: [] for k in ("row_type", "col_type", "row_l", "col_l", "bond", "kind", "off", "n",
                                "r_cut", "lam", "r_on")}
        rx, ry, ry2 = [], [], []
        for (rt, ct), p in sorted(self.pairs.items()):
            pair_cut[rt * nt + ct] = p.s_max_cutoff if which == "S" else p.h_max_cutoff
            for c in p.channels:
                rad = c.s if which == "S" else c.h
                if rad is None:
                    continue
                cols["row_type"].append(rt)
                cols["col_type"].append(ct)
                cols["row_l"].append(c.row_l)
                cols["col_l"].append(c.col_l)
                cols["bond"].append(c.channel)
                off = sum(len(a) for a in ry)
                cols["off"].append(off)
                if rad.kind == "spline":
                    cols["kind"].append(RADIAL_SPLINE)
                    cols["n"].append(rad.r.size)
                    rx.append(rad.r)
                    ry.append(rad.values)
                    ry2.append(rad.y2)
                else:
                    cols["kind"].append(RADIAL_EXP_POLY)
                    cols["n"].append(rad.coeffs.size)
                    rx.append(np.zeros(rad.coeffs.size))
                    ry.append(rad.coeffs)
                    ry2.append(np.zeros(rad.coeffs.size))
                cols["r_cut"].append(rad.r_cut)
                cols["lam"].append(rad.lambda_value)
                cols["r_on"].append(rad.r_on)
        out = {
            "n_types": nt, "norb": b,
            "orb_l": orb_l, "orb_m": orb_m, "onsite": np.ascontiguousarray(onsite, dtype=np.float64),
            "pair_cutoff": pair_cut,
            "rad_x": np.concatenate(rx) if rx else np.zeros(1),
            "rad_y": np.concatenate(ry) if ry else np.zeros(1),
            "rad_y2": np.concatenate(ry2) if ry2 else np.zeros(1),
        }
        for k in ("row_type", "col_type", "row_l", "col_l", "bond", "kind", "off", "n"):
            out["ch_" + k] = np.array(cols[k], dtype=np.int32)
        for k in ("r_cut", "lam", "r_on"):
            out["ch_" + k] = np.array(cols[k], dtype=np.float64)
        out["n_channels"] = len(cols["row_type"])
        return out


class SktbModel:
    """Slater-Koster tight-binding model whose H and S are assembled on the
    device (SktbModel of the reference, rsatb/model/sktb/sktb_model.hpp, over
    SkTableConstruction). Structure atom types index the table's types, which
    are ordered by atomic number (sk_table.hpp:289-298)."""

    device_assembly = True

    def __init__(self, basis: dict, hamiltonian: SkParameters, overlap: SkParameters | None = None):
        self.table = SkTable(basis, hamiltonian, overlap)
        norb = self.table.type_norb
        if len(set(norb)) != 1:
            raise ValueError("the device BSR holds uniform blocks: every atom type needs the same orbital count")
        self.norb = norb[0]
        self._arrays = {"H": self.table.arrays("H")}
        if self.table.has_overlap:
            self._arrays["S"] = self.table.arrays("S")

    @property
    def cutoff(self) -> float:
        """Neighbour-list cutoff: the largest bond cutoff of any pair (each
        pair's r_i + r_j of sk_table.hpp:176-191 is at most this)."""
        return max((max(p.h_max_cutoff, p.s_max_cutoff) for p in self.table.pairs.values()), default=0.0)

    def has_overlap(self) -> bool:
        return self.table.has_overlap

    def arrays(self, which: str) -> dict:
        return self._arrays[which]
