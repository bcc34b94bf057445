"""rsatb_amd: MI355X (gfx950) stochastic Chebyshev/KPM path of RSATB.

The compute path is librsatb_gpu.so (hand-written HIP kernels behind the C ABI
of include/rsatb_gpu.h); this package assembles operands and mirrors rsatb's
Python API (RSATB, SetupOptions, PreconditionerOptions).
"""
from .bsr import BSR
from .driver import RSATB, PreconditionerOptions, SetupOptions, make_explicit_bounds
from .lattice import (Structure, diamond_supercell, graphene_primitive_cell, graphene_rect_supercell,
                      neighbor_edges)
from .models import RandomBlockModel, SkPzModel, assemble

__all__ = [
    "BSR", "RSATB", "PreconditionerOptions", "SetupOptions", "make_explicit_bounds", "Structure",
    "diamond_supercell", "graphene_primitive_cell", "graphene_rect_supercell", "neighbor_edges",
    "RandomBlockModel", "SkPzModel", "assemble",
]
