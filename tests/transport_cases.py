"""A small graphene p_z transport case with a linear hopping
parameterization (onsite e0, three hopping shells), shared by the CPU and
GPU conductivity-derivative tests."""
from __future__ import annotations

import numpy as np

from rsatb_amd.autograd import linear_transport_model
from rsatb_amd.lattice import graphene_rect_supercell, neighbor_edges

RADII = (1.6, 2.5, 2.9)
THETA = np.array([0.3, -2.7, -0.2, -0.18])


def graphene_transport(nx: int = 4, ny: int = 4):
    """(structure, edges, model, edge_coef, onsite_coef): parameter 0 is the
    onsite energy, parameter k >= 1 the hopping of shell k (first radius
    that holds the bond)."""
    st = graphene_rect_supercell(nx, ny)
    e = neighbor_edges(st, RADII[-1])
    shell = np.full(e.src.size, -1)
    for k, r in sorted(enumerate(RADII), key=lambda t: t[1], reverse=True):
        shell[e.dist <= r] = k
    ec = np.zeros((e.src.size, 1 + len(RADII)))
    for k in range(len(RADII)):
        ec[shell == k, k + 1] = 1.0
    oc = np.zeros((st.n_atoms, 1 + len(RADII)))
    oc[:, 0] = 1.0
    model = linear_transport_model(st.n_atoms, e, ec, oc, ("e0", "t1", "t2", "t3"))
    return st, e, model, ec, oc
