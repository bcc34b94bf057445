"""Shared Slater-Koster cases of the SK tests (CPU oracle and GPU parity)."""
from __future__ import annotations

import numpy as np

from rsatb_amd.sktb import Radial, SkParameters, SkTable


def silicon_params(exp_poly=True):
    """(basis, H, S) of an sp3d5 silicon-like model with every channel;
    radial kinds mixed."""
    h = SkParameters()
    h.onsite["Si"] = {"s": -4.2, "p": 1.7, "d": 6.5}
    vals = {("ss", "sigma"): -1.9, ("sp", "sigma"): 2.1, ("pp", "sigma"): 3.0, ("pp", "pi"): -0.9,
            ("sd", "sigma"): -1.6, ("pd", "sigma"): -1.3, ("pd", "pi"): 1.9, ("dd", "sigma"): -1.2,
            ("dd", "pi"): 2.2, ("dd", "delta"): -1.7}
    for i, ((sp, bond), v) in enumerate(vals.items()):
        if exp_poly and i % 2:
            rad = Radial.exp_poly([v, 0.1 * v, -0.05 * v], 3.2, 0.4, 2.6)
        else:
            r = np.linspace(1.8, 3.2, 8)
            rad = Radial.spline(r, v * np.exp(-(r - 2.35)) * (1 + 0.1 * np.cos(3 * r)), 3.2)
        h.set_pair_channel("Si", "Si", sp, bond, rad)
    s = SkParameters()
    for i, ((sp, bond), v) in enumerate(vals.items()):
        r = np.linspace(1.8, 3.2, 6)
        s.set_pair_channel("Si", "Si", sp, bond, Radial.spline(r, 0.03 * v * np.exp(-(r - 2.35)), 3.2))
    return {"Si": ["s", "p", "d"]}, h, s


def sp3_silicon_table(exp_poly=True):
    return SkTable(*silicon_params(exp_poly))


def periodic_edges(pos, cell, cutoff):
    """All (i, j, R) with 0 < |pos_j + R cell - pos_i| <= cutoff, R in {-1,0,1}^3."""
    src, dst, sh = [], [], []
    n = len(pos)
    for R in np.array(np.meshgrid([-1, 0, 1], [-1, 0, 1], [-1, 0, 1])).T.reshape(-1, 3):
        d = pos[None, :, :] + (R @ cell)[None, None, :] - pos[:, None, :]
        dist = np.linalg.norm(d, axis=2)
        ii, jj = np.nonzero((dist > 1e-9) & (dist <= cutoff))
        src.append(ii)
        dst.append(jj)
        sh.append(np.tile(R, (ii.size, 1)))
    return np.stack([np.concatenate(src), np.concatenate(dst)]), np.concatenate(sh)


def silicon_cell():
    a = 5.43
    base = np.array([[0, 0, 0], [0.25, 0.25, 0.25]])
    fcc = np.array([[0, 0, 0], [0, 0.5, 0.5], [0.5, 0, 0.5], [0.5, 0.5, 0]])
    pos = np.concatenate([fcc + b for b in base]) * a
    return pos, np.eye(3) * a
