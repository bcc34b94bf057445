"""Structures and neighbour lists for the synthetic systems the path runs on.

Conventions followed (paths relative to the RSATB source tree):
  - edge i -> j through lattice shift R iff |r_j + R.cell - r_i| <= cutoff
    (rsatb/model/basis_spec.hpp:83-86 with cutoff = r_i + r_j);
  - bond vector r_j + R.cell - r_i (rsatb/rsatb.hpp:66-69);
  - every bond appears twice, as (i, j, R) and (j, i, -R) (from_graph
    convention, rsatb/driver.py:73-115, test/test_graphene_pz_dos.py:68-75);
  - canonical order: edges sorted by (i, j, R); rows are atoms in input order.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from itertools import product

import numpy as np

try:  # scipy is part of the image; the brute-force path covers its absence
    from scipy.spatial import cKDTree
except Exception:  # pragma: no cover
    cKDTree = None


@dataclass
class Structure:
    positions: np.ndarray            # [N, 3] Cartesian, Angstrom
    cell: np.ndarray                 # [3, 3], rows are lattice vectors
    pbc: tuple = (True, True, True)
    atom_type: np.ndarray = field(default=None)  # [N] int

    def __post_init__(self):
        self.positions = np.ascontiguousarray(self.positions, dtype=np.float64)
        self.cell = np.ascontiguousarray(self.cell, dtype=np.float64).reshape(3, 3)
        if self.atom_type is None:
            self.atom_type = np.zeros(len(self.positions), dtype=np.int32)
        self.atom_type = np.ascontiguousarray(self.atom_type, dtype=np.int32)

    @property
    def n_atoms(self) -> int:
        return int(self.positions.shape[0])


@dataclass
class Edges:
    src: np.ndarray    # [E] int64
    dst: np.ndarray    # [E] int64
    shift: np.ndarray  # [E, 3] int64 lattice shift of dst
    vec: np.ndarray    # [E, 3] bond vector r_dst + shift.cell - r_src
    dist: np.ndarray   # [E]


def tile(unit: Structure, reps: tuple) -> Structure:
    """Supercell: unit cell repeated reps = (n1, n2, n3) times; atoms ordered
    by (i1, i2, i3, basis)."""
    n1, n2, n3 = reps
    idx = np.array(list(product(range(n1), range(n2), range(n3))), dtype=np.float64)
    shifts = idx @ unit.cell
    pos = (shifts[:, None, :] + unit.positions[None, :, :]).reshape(-1, 3)
    types = np.tile(unit.atom_type, len(idx))
    cell = unit.cell * np.array([[n1], [n2], [n3]], dtype=np.float64)
    return Structure(pos, cell, unit.pbc, types)


def graphene_rect_cell(a_cc: float = 1.42, vacuum: float = 20.0) -> Structure:
    """4-atom rectangular honeycomb cell: A1 = (sqrt3 a, 0, 0), A2 = (0, 3 a, 0)."""
    s3 = np.sqrt(3.0)
    cell = np.array([[s3 * a_cc, 0, 0], [0, 3 * a_cc, 0], [0, 0, vacuum]])
    z = 0.5 * vacuum
    pos = np.array([
        [0.0, 0.0, z],
        [0.0, a_cc, z],
        [0.5 * s3 * a_cc, 1.5 * a_cc, z],
        [0.5 * s3 * a_cc, 2.5 * a_cc, z],
    ])
    return Structure(pos, cell, (True, True, False))


def graphene_rect_supercell(nx: int, ny: int, a_cc: float = 1.42, vacuum: float = 20.0) -> Structure:
    return tile(graphene_rect_cell(a_cc, vacuum), (nx, ny, 1))


def graphene_primitive_cell(a_cc: float = 1.42, vacuum: float = 20.0) -> Structure:
    """2-atom cell of test/test_graphene_pz_dos.py:22-33."""
    a = np.sqrt(3.0) * a_cc
    cell = np.array([[a, 0, 0], [0.5 * a, 1.5 * a_cc, 0], [0, 0, vacuum]])
    pos = np.array([[0.0, 0.0, 0.5 * vacuum], [0.0, a_cc, 0.5 * vacuum]])
    return Structure(pos, cell, (True, True, False))


def twisted_bilayer_graphene(m: int = 31, n: int = 32, a_cc: float = 1.42, layer_gap: float = 3.35,
                             vacuum: float = 20.0) -> Structure:
    """Commensurate twisted bilayer graphene cell of index (m, n): superlattice
    vectors T1 = m a1 + n a2 and T2 = -n a1 + (m + n) a2 of the bottom layer,
    the top layer rotated by the angle phi that takes n a1 + m a2 onto T1
    (cos phi = (m^2 + 4 m n + n^2) / (2 (m^2 + m n + n^2)); (31, 32) gives
    1.05 degrees and 2 (m^2 + m n + n^2) = 5954 atoms per layer). Bottom
    layer atoms first; atom_type 0 bottom, 1 top."""
    a = np.sqrt(3.0) * a_cc
    a1 = np.array([a, 0.0])
    a2 = np.array([0.5 * a, 0.5 * np.sqrt(3.0) * a])
    basis = [np.zeros(2), (a1 + a2) / 3.0]
    T = np.stack([m * a1 + n * a2, -n * a1 + (m + n) * a2])
    Tinv = np.linalg.inv(T)
    u = n * a1 + m * a2
    phi = np.arctan2(u[0] * T[0, 1] - u[1] * T[0, 0], u @ T[0])
    rot = np.array([[np.cos(phi), -np.sin(phi)], [np.sin(phi), np.cos(phi)]])
    n_layer = 2 * (m * m + m * n + n * n)
    span = 2 * (m + n) + 2
    p, q = np.meshgrid(np.arange(-span, span + 1), np.arange(-span, span + 1), indexing="ij")
    lat = p.reshape(-1, 1) * a1 + q.reshape(-1, 1) * a2
    layers = []
    for R in (np.eye(2), rot):
        pts = np.concatenate([(lat + b) @ R.T for b in basis])
        f = pts @ Tinv
        f = f - np.floor(f + 1e-9)  # wrap into [0, 1) with a tolerance at the edges
        key = np.round(f * 1e6).astype(np.int64)
        _, first = np.unique(key, axis=0, return_index=True)
        f = f[np.sort(first)]
        if f.shape[0] != n_layer:
            raise RuntimeError(f"twisted cell ({m}, {n}): {f.shape[0]} atoms in a layer, expected {n_layer}")
        layers.append(f @ T)
    z0 = 0.5 * vacuum
    pos = np.concatenate([np.column_stack([layers[0], np.full(n_layer, z0)]),
                          np.column_stack([layers[1], np.full(n_layer, z0 + layer_gap)])])
    cell = np.zeros((3, 3))
    cell[:2, :2] = T
    cell[2, 2] = vacuum + layer_gap
    types = np.concatenate([np.zeros(n_layer, np.int32), np.ones(n_layer, np.int32)])
    return Structure(pos, cell, (True, True, False), types)


def diamond_cell(a: float = 3.567) -> Structure:
    """8-atom conventional diamond cell."""
    fcc = np.array([[0, 0, 0], [0, 0.5, 0.5], [0.5, 0, 0.5], [0.5, 0.5, 0]])
    frac = np.concatenate([fcc, fcc + 0.25])
    return Structure(frac * a, np.eye(3) * a, (True, True, True))


def diamond_supercell(n: int, a: float = 3.567) -> Structure:
    return tile(diamond_cell(a), (n, n, n))


def neighbor_edges(st: Structure, cutoff: float) -> Edges:
    """All directed edges within `cutoff` (self at zero shift excluded)."""
    pos, cell = st.positions, st.cell
    inv = np.linalg.inv(cell)
    frac = pos @ inv
    # two points within `cutoff` differ in fractional coordinate d by at most
    # c_d = cutoff * |inv[:, d]|; only shifts that can bring some image that
    # close are visited
    fmin, fmax = frac.min(0), frac.max(0)
    c = cutoff * np.linalg.norm(inv, axis=0)
    ranges = []
    for d in range(3):
        if st.pbc[d]:
            lo = int(np.floor(fmin[d] - fmax[d] - c[d]))
            hi = int(np.ceil(fmax[d] - fmin[d] + c[d]))
            ranges.append(range(lo, hi + 1))
        else:
            ranges.append(range(0, 1))
    src_l, dst_l, sh_l = [], [], []
    tree = cKDTree(pos) if cKDTree is not None else None
    for sh in product(*ranges):
        sv = np.asarray(sh, dtype=np.float64)
        if np.any(fmin + sv - fmax > c) or np.any(fmax + sv - fmin < -c):
            continue
        img = pos + sv @ cell
        if tree is not None:
            other = cKDTree(img)
            pairs = tree.sparse_distance_matrix(other, cutoff, output_type="ndarray")
            i = pairs["i"].astype(np.int64)
            j = pairs["j"].astype(np.int64)
        else:
            d = np.linalg.norm(img[None, :, :] - pos[:, None, :], axis=-1)
            i, j = np.nonzero(d <= cutoff)
        if sh == (0, 0, 0):
            keep = i != j
            i, j = i[keep], j[keep]
        if i.size == 0:
            continue
        src_l.append(i)
        dst_l.append(j)
        sh_l.append(np.broadcast_to(np.asarray(sh, dtype=np.int64), (i.size, 3)))
    if src_l:
        src = np.concatenate(src_l)
        dst = np.concatenate(dst_l)
        shift = np.concatenate(sh_l)
    else:
        src = dst = np.zeros(0, dtype=np.int64)
        shift = np.zeros((0, 3), dtype=np.int64)
    vec = pos[dst] + shift @ cell - pos[src]
    dist = np.linalg.norm(vec, axis=1)
    keep = dist <= cutoff
    src, dst, shift, vec, dist = src[keep], dst[keep], shift[keep], vec[keep], dist[keep]
    order = np.lexsort((shift[:, 2], shift[:, 1], shift[:, 0], dst, src))
    return Edges(src[order], dst[order], shift[order], vec[order], dist[order])


def edges_from_graph(st: Structure, edge_index, edge_shift_vec) -> Edges:
    """Edges given as from_graph arrays: flattened (src, dst) pairs and
    flattened integer shifts (rsatb/driver.py:73-115)."""
    ei = np.asarray(edge_index, dtype=np.int64).reshape(-1, 2)
    sh = np.asarray(edge_shift_vec, dtype=np.int64).reshape(-1, 3)
    if ei.shape[0] != sh.shape[0]:
        raise ValueError("edge_index and edge_shift_vec describe different edge counts")
    src, dst = ei[:, 0], ei[:, 1]
    vec = st.positions[dst] + sh @ st.cell - st.positions[src]
    return Edges(src, dst, sh, vec, np.linalg.norm(vec, axis=1))
