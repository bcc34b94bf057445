"""CPU restatement of the reference's Slater-Koster block construction
(TEST INFRASTRUCTURE ONLY: tests/, and bench timing of the CPU path, use it
as the checker; the product never imports it).

Follows rsatb/model/sktb/sktb_construction.hpp and radial.hpp:
  - radial_eval: spline (:139-172, natural cubic with the second derivatives
    of :93-125, constant below the first knot, last value between the last
    knot and r_cut, zero beyond r_cut) and exp_poly (:174-195, cosine taper
    0.5 (1 + cos(pi (d - r_on) / (r_cut - r_on))) beyond r_on).
  - local_frame (:117-131): ez along the bond, ex = normalise(z x ez) unless
    the bond is along z, ey = ez x ex; columns (ex, ey, ez).
  - shell rotations (:140-177): p rows/cols in magnetic order (py, pz, px);
    d through the Frobenius products of the canonical real quadrupoles
    rotated by the frame.
  - aligned_shell_block (:222-244): sigma on (m = 0, m = 0), pi/delta on
    (+-m, +-m), sign -1 when row_l > col_l and row_l + col_l is odd.
  - rotate_shell_block (:246-275): R_row^T A R_col.
  - pair_block (:437-500): direct table pair if present and within its
    cutoff, else the transpose of the reversed pair at -displacement.
  - assemble (:327-425): onsite diagonal block for every atom, one block per
    edge unless empty or all |v| <= 1e-14, images summed at the k-point with
    phase exp(2 pi i k.R) (the host assembler's convention).
Pinned in tests/test_sktb.py against the known answers of
rsatb/model/sktb/test/test_sktb.cpp and closed forms (direction-cosine p-p,
quadrupole d-d and p-d outer products, orthogonal rotations).
"""
from __future__ import annotations

import bisect
import math

import numpy as np

_S2 = 1.0 / math.sqrt(2.0)
_S6 = 1.0 / math.sqrt(6.0)
# canonical real d quadrupoles in magnetic order dxy, dyz, dz2, dzx, dx2-y2 (:93-115)
CANON_D = np.array([
    [[0, _S2, 0], [_S2, 0, 0], [0, 0, 0]],
    [[0, 0, 0], [0, 0, _S2], [0, _S2, 0]],
    [[-_S6, 0, 0], [0, -_S6, 0], [0, 0, 2 * _S6]],
    [[0, 0, _S2], [0, 0, 0], [_S2, 0, 0]],
    [[_S2, 0, 0], [0, -_S2, 0], [0, 0, 0]],
])
_P_AXIS = (1, 2, 0)  # py, pz, px (:135-138)


def radial_eval(rad, d: float) -> float:
    if d < 0.0:
        return 0.0
    if rad.kind == "spline":
        if d > rad.r_cut:
            return 0.0
        r, v, y2 = rad.r, rad.values, rad.y2
        if d <= r[0]:
            return float(v[0])
        i = bisect.bisect_left(list(r), d)
        if i == len(r):
            return float(v[-1])
        lo, hi = i - 1, i
        h = r[hi] - r[lo]
        if h <= 1e-12:
            return float(v[hi])
        a = (r[hi] - d) / h
        b = (d - r[lo]) / h
        return float(a * v[lo] + b * v[hi] + ((a * a * a - a) * y2[lo] + (b * b * b - b) * y2[hi]) * (h * h) / 6.0)
    if rad.kind == "exp_poly":
        if d > rad.r_cut:
            return 0.0
        w = 1.0
        if d > rad.r_on:
            span = rad.r_cut - rad.r_on
            w = 0.0 if span <= 1e-12 else 0.5 * (1.0 + math.cos(math.pi * (d - rad.r_on) / span))
        poly, power = 0.0, 1.0
        for c in rad.coeffs:
            poly += c * power
            power *= d
        return w * math.exp(-rad.lambda_value * d) * poly
    raise ValueError(f"unsupported radial kind {rad.kind}")


def _normalized(v):
    n = math.sqrt(float(v @ v))
    return np.zeros(3) if n <= 1e-12 else v / n


def local_frame(direction) -> np.ndarray:
    ez = _normalized(np.asarray(direction, dtype=np.float64))
    ex = np.array([1.0, 0.0, 0.0])
    if abs(ez[2]) < 1.0 - 1e-12:
        ex = _normalized(np.cross([0.0, 0.0, 1.0], ez))
    ey = np.cross(ez, ex)
    return np.stack([ex, ey, ez], axis=1)


def shell_rotation(l: int, frame: np.ndarray) -> np.ndarray:
    """rotation[local, global] (:140-177)."""
    if l == 0:
        return np.ones((1, 1))
    if l == 1:
        R = np.zeros((3, 3))
        for loc in range(3):
            for glo in range(3):
                R[loc, glo] = frame[_P_AXIS[glo], _P_AXIS[loc]]
        return R
    if l == 2:
        R = np.zeros((5, 5))
        for loc in range(5):
            rot = frame @ CANON_D[loc] @ frame.T
            for glo in range(5):
                R[loc, glo] = float(np.sum(CANON_D[glo] * rot))
        return R
    raise ValueError("SKTB shell rotation supports only s, p, and d orbitals.")


def aligned_shell_block(row_l: int, col_l: int, channel: int, value: float) -> np.ndarray:
    blk = np.zeros((2 * row_l + 1, 2 * col_l + 1))
    sign = -1.0 if (row_l > col_l and (row_l + col_l) % 2 != 0) else 1.0
    v = sign * value
    if channel == 0:
        blk[row_l, col_l] = v
    else:
        blk[row_l - channel, col_l - channel] = v
        blk[row_l + channel, col_l + channel] = v
    return blk


def _direct_block(table, which: str, rt: int, ct: int, vec):
    pair = table.pairs.get((rt, ct))
    if pair is None:
        return None
    d = math.sqrt(float(np.dot(vec, vec)))
    cut = pair.s_max_cutoff if which == "S" else pair.h_max_cutoff
    if d <= 1e-12 or d > cut + 1e-12:
        return None
    frame = local_frame(vec)
    rots = [shell_rotation(l, frame) for l in range(3)]
    ra, ca = table.atom_types[rt], table.atom_types[ct]
    blk = np.zeros((len(ra.orbitals), len(ca.orbitals)))
    for ch in pair.channels:
        rad = ch.s if which == "S" else ch.h
        if rad is None or d > rad.r_cut + 1e-12:
            continue
        aligned = aligned_shell_block(ch.row_l, ch.col_l, ch.channel, radial_eval(rad, d))
        rotated = rots[ch.row_l].T @ aligned @ rots[ch.col_l]
        for i, (_, li, mi) in enumerate(ra.orbitals):
            if li != ch.row_l:
                continue
            for j, (_, lj, mj) in enumerate(ca.orbitals):
                if lj == ch.col_l:
                    blk[i, j] += rotated[mi, mj]
    return blk


def pair_block(table, which: str, rt: int, ct: int, vec):
    """Row-major block of an edge from an atom of type rt to one of type ct
    along `vec`, or None when neither ordering of the pair applies."""
    vec = np.asarray(vec, dtype=np.float64)
    blk = _direct_block(table, which, rt, ct, vec)
    if blk is not None:
        return blk
    rev = _direct_block(table, which, ct, rt, -vec)
    return None if rev is None else rev.T.copy()


def edge_vectors(pos, cell, edge_index, edge_shift) -> np.ndarray:
    """pos[j] + R . cell - pos[i] with the cell rows as lattice vectors."""
    pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
    cell = np.asarray(cell, dtype=np.float64).reshape(3, 3)
    ei = np.asarray(edge_index).reshape(2, -1)
    sh = np.asarray(edge_shift, dtype=np.float64).reshape(-1, 3)
    return pos[ei[1]] + sh @ cell - pos[ei[0]]


def assemble(table, which: str, atom_type, pos, cell, edge_index, edge_shift, kpoint=(0.0, 0.0, 0.0)):
    """(row_ptr, col_ind, blocks row-major [nnzb, b, b]) of H(k) or S(k)."""
    atom_type = np.asarray(atom_type)
    n = atom_type.size
    b = table.type_norb[0]
    ei = np.asarray(edge_index).reshape(2, -1)
    shifts = np.asarray(edge_shift, dtype=np.int64).reshape(-1, 3)
    vecs = edge_vectors(pos, cell, ei, shifts)
    k = np.asarray(kpoint, dtype=np.float64)
    gamma = not np.any(k)
    dt = np.float64 if gamma else np.complex128
    acc: dict[tuple[int, int], np.ndarray] = {}
    for a in range(n):
        t = table.atom_types[atom_type[a]]
        acc[(a, a)] = np.diag(t.s_onsite if which == "S" else t.h_onsite).astype(dt)
    for e in range(ei.shape[1]):
        i, j = int(ei[0, e]), int(ei[1, e])
        blk = pair_block(table, which, int(atom_type[i]), int(atom_type[j]), vecs[e])
        if blk is None or not np.any(np.abs(blk) > 1e-14):
            continue
        if not gamma:
            blk = blk * np.exp(2j * np.pi * float(shifts[e] @ k))
        if (i, j) in acc:
            acc[(i, j)] = acc[(i, j)] + blk
        else:
            acc[(i, j)] = blk.astype(dt)
    keys = sorted(acc)
    row_ptr = np.zeros(n + 1, dtype=np.int64)
    for (r, _) in keys:
        row_ptr[r + 1] += 1
    row_ptr = np.cumsum(row_ptr)
    col_ind = np.array([c for (_, c) in keys], dtype=np.int32)
    blocks = np.stack([acc[kk] for kk in keys]) if keys else np.zeros((0, b, b), dt)
    return row_ptr, col_ind, blocks
