"""Tight-binding models that produce the BSR operands of the path.

Two families cover the workloads:
  - SkPzModel: one p_z orbital per atom with Slater-Koster two-centre
    hoppings V(d) = n^2 V_pp_sigma(d) + (1 - n^2) V_pp_pi(d), n = z / d (the
    pz-pz Slater-Koster rule), tabulated as distance shells, plus an optional
    overlap of the same form. Onsite blocks are stored for every atom even
    when zero and all-zero edge blocks are dropped, as rsatb's SK assembly
    does (rsatb/model/sktb/sktb_construction.hpp:338-360, 385).
  - RandomBlockModel: b orbitals per atom with random symmetric blocks,
    H_ij = N(0, 1) e^{-r/decay} and S = I + N(0, s_sigma) e^{-r/decay} off the
    diagonal, drawn per unordered bond from a counter-based generator so the
    matrices do not depend on edge order; H_ji = H_ij^T, S likewise.

Assembly (`assemble`) sums the images of every pair into one block at the
requested k-point, H(k)_ij = sum_R H_ij(R) e^{2 pi i k.R} with R the integer
lattice shift of the edge; at Gamma the operands stay real.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .bsr import BSR
from .lattice import Edges, Structure

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def mix64(x: np.ndarray) -> np.ndarray:
    """SplitMix64 finaliser on uint64 arrays (wrapping arithmetic)."""
    with np.errstate(over="ignore"):
        x = (x + np.uint64(0x9E3779B97F4A7C15)) & _M64
        x = ((x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _M64
        x = ((x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _M64
        return x ^ (x >> np.uint64(31))


def _uniform(h: np.ndarray) -> np.ndarray:
    """(0, 1) doubles from the top 53 bits."""
    return ((h >> np.uint64(11)).astype(np.float64) + 0.5) * (1.0 / 9007199254740992.0)


def counter_normal(keys: np.ndarray, n: int) -> np.ndarray:
    """n standard normals per uint64 key (Box-Muller on SplitMix64 draws)."""
    keys = np.asarray(keys, dtype=np.uint64)
    e = np.arange(n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        base = keys[:, None] ^ (e[None, :] * np.uint64(0x632BE59BD9B4E019))
    u1 = _uniform(mix64(base * np.uint64(2) & _M64))
    u2 = _uniform(mix64((base * np.uint64(2) + np.uint64(1)) & _M64))
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


def bond_keys(seed: int, tag: int, i: np.ndarray, j: np.ndarray, shift: np.ndarray) -> np.ndarray:
    s = np.asarray(shift, dtype=np.int64) & 0xFFFF
    packed = (s[:, 0] | (s[:, 1] << 16) | (s[:, 2] << 32)).astype(np.uint64)
    h = mix64(np.full(i.shape, np.uint64(seed) ^ np.uint64(tag), dtype=np.uint64))
    h = mix64(h ^ i.astype(np.uint64))
    h = mix64(h ^ j.astype(np.uint64))
    return mix64(h ^ packed)


def canonical_bond(edges: Edges) -> np.ndarray:
    """True for the edge of each (i, j, R) / (j, i, -R) pair that draws the
    block: i < j, or i == j with R lexicographically positive."""
    s = edges.shift
    pos_shift = (s[:, 0] > 0) | ((s[:, 0] == 0) & ((s[:, 1] > 0) | ((s[:, 1] == 0) & (s[:, 2] > 0))))
    return (edges.src < edges.dst) | ((edges.src == edges.dst) & pos_shift)


@dataclass
class SkPzModel:
    """p_z tight binding. Shells are (r_max, V_pp_pi, V_pp_sigma); the first
    shell whose r_max >= d applies, beyond the last one the hopping is zero."""

    onsite: float = 0.0
    shells: list = field(default_factory=lambda: [(1.6, -2.7, 0.0)])
    overlap_shells: list | None = None

    norb: int = 1

    @property
    def cutoff(self) -> float:
        r = max(s[0] for s in self.shells)
        if self.overlap_shells:
            r = max(r, max(s[0] for s in self.overlap_shells))
        return r

    @staticmethod
    def _shell_value(shells, d: np.ndarray, n2: np.ndarray) -> np.ndarray:
        out = np.zeros_like(d)
        done = np.zeros(d.shape, dtype=bool)
        for r_max, v_pi, v_sigma in sorted(shells, key=lambda t: t[0]):
            sel = (~done) & (d <= r_max)
            out[sel] = n2[sel] * v_sigma + (1.0 - n2[sel]) * v_pi
            done |= sel
        return out

    def blocks(self, st: Structure, edges: Edges, which: str):
        n2 = (edges.vec[:, 2] / np.maximum(edges.dist, 1e-300)) ** 2
        if which == "H":
            onsite = np.full((st.n_atoms, 1, 1), float(self.onsite))
            hop = self._shell_value(self.shells, edges.dist, n2)
        else:
            if not self.overlap_shells:
                return None
            onsite = np.ones((st.n_atoms, 1, 1))
            hop = self._shell_value(self.overlap_shells, edges.dist, n2)
        return onsite, hop[:, None, None]

    def has_overlap(self) -> bool:
        return bool(self.overlap_shells)


@dataclass
class RandomBlockModel:
    """LCAO-like random blocks (b orbitals per atom) within `cutoff`."""

    norb: int = 13
    cutoff: float = 2.7
    decay: float = 1.5
    h_sigma: float = 1.0
    s_sigma: float = 0.05
    seed: int = 20240917
    overlap: bool = True

    def has_overlap(self) -> bool:
        return self.overlap

    def blocks(self, st: Structure, edges: Edges, which: str):
        b = self.norb
        if which == "S" and not self.overlap:
            return None
        tag = 1 if which == "H" else 2
        canon = canonical_bond(edges)
        ci = np.flatnonzero(canon)
        keys = bond_keys(self.seed, tag, edges.src[ci], edges.dst[ci], edges.shift[ci])
        sigma = self.h_sigma if which == "H" else self.s_sigma
        draw = counter_normal(keys, b * b).reshape(-1, b, b)
        draw *= (sigma * np.exp(-edges.dist[ci] / self.decay))[:, None, None]
        hop = np.zeros((edges.src.size, b, b))
        hop[ci] = draw
        # the reverse edge (j, i, -R) carries the transpose
        nc = np.flatnonzero(~canon)
        if nc.size:
            rev = _reverse_index(edges)
            hop[nc] = np.transpose(hop[rev[nc]], (0, 2, 1))
        if which == "H":
            okeys = bond_keys(self.seed, 3, np.arange(st.n_atoms), np.arange(st.n_atoms),
                              np.zeros((st.n_atoms, 3), dtype=np.int64))
            m = counter_normal(okeys, b * b).reshape(-1, b, b) * self.h_sigma
            onsite = np.triu(m) + np.transpose(np.triu(m, 1), (0, 2, 1))
        else:
            onsite = np.broadcast_to(np.eye(b), (st.n_atoms, b, b)).copy()
        return onsite, hop


def _reverse_index(edges: Edges) -> np.ndarray:
    """For every edge the index of its reverse (j, i, -R)."""
    n = edges.src.size
    key_f = np.stack([edges.src, edges.dst, edges.shift[:, 0], edges.shift[:, 1], edges.shift[:, 2]], 1)
    key_r = np.stack([edges.dst, edges.src, -edges.shift[:, 0], -edges.shift[:, 1], -edges.shift[:, 2]], 1)
    order_f = np.lexsort(key_f.T[::-1])
    order_r = np.lexsort(key_r.T[::-1])
    if not np.array_equal(key_f[order_f], key_r[order_r]):
        raise ValueError("edge list is not symmetric: every (i, j, R) needs its (j, i, -R)")
    rev = np.empty(n, dtype=np.int64)
    rev[order_r] = order_f
    return rev


def assemble(st: Structure, edges: Edges, model, which: str = "H", kpoint=(0.0, 0.0, 0.0)) -> BSR | None:
    """BSR of H(k) or S(k) with onsite blocks stored for every atom and
    all-zero edge blocks dropped."""
    blk = model.blocks(st, edges, which)
    if blk is None:
        return None
    onsite, hop = blk
    b = onsite.shape[1]
    k = np.asarray(kpoint, dtype=np.float64)
    gamma = not np.any(k)
    nz = np.any(hop.reshape(hop.shape[0], -1) != 0.0, axis=1)
    src, dst, sh, hop = edges.src[nz], edges.dst[nz], edges.shift[nz], hop[nz]
    if not gamma:
        phase = np.exp(2j * np.pi * (sh @ k))
        hop = hop * phase[:, None, None]
        onsite = onsite.astype(np.complex128)
    n = st.n_atoms
    rows = np.concatenate([np.arange(n, dtype=np.int64), src])
    cols = np.concatenate([np.arange(n, dtype=np.int64), dst])
    blocks = np.concatenate([onsite, hop], axis=0)
    return BSR.from_blocks(n, b, rows, cols, blocks)


def random_block_reverse_check(edges: Edges) -> np.ndarray:
    """Exposes the reverse-edge map (tests use it to check symmetry)."""
    return _reverse_index(edges)
