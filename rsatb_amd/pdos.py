"""Row groupings of the projected DOS (rsatb/calc/pdos.hpp:56-127).

A grouping maps every scalar row (orbital) of the operator to a group id in
[0, n_groups) or to -1 (no group). The moments themselves come from the
library (rg_kpm_pdos_moments); this module only builds the map, in the
reference's two flavours:

  by atom           one group per atom, id = the atom's input index
                    (PdosGroupsByAtom, pdos.hpp:72-88);
  by orbital slot   one group per (atom type, orbital offset) pair, types in
                    ascending order of their ids, so two species never share a
                    group (PdosGroupsByOrbitalSlot, pdos.hpp:90-127).

Rows are atom-major: atom a owns type_norb[atom_type[a]] consecutive rows.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass
class PdosGroups:
    n_groups: int
    group_of_row: np.ndarray                      # int32 [rows]
    group_type: np.ndarray = field(default=None)  # [n_groups] by orbital slot, else empty
    group_slot: np.ndarray = field(default=None)

    def __post_init__(self):
        self.group_of_row = np.ascontiguousarray(self.group_of_row, dtype=np.int32)
        if self.group_type is None:
            self.group_type = np.zeros(0, dtype=np.int32)
        if self.group_slot is None:
            self.group_slot = np.zeros(0, dtype=np.int32)


def _norb_per_atom(atom_type, type_norb) -> np.ndarray:
    at = np.asarray(atom_type, dtype=np.int64).reshape(-1)
    tn = np.asarray(type_norb, dtype=np.int64).reshape(-1)
    if at.size and (at.min() < 0 or at.max() >= tn.size):
        raise RuntimeError("atom_type out of range of type_norb")
    return tn[at]


def pdos_groups_by_atom(atom_type, type_norb, atom_index=None) -> PdosGroups:
    """One group per atom; id = atom_index[a] (default: input order)."""
    norb = _norb_per_atom(atom_type, type_norb)
    idx = np.arange(norb.size) if atom_index is None else np.asarray(atom_index, dtype=np.int64).reshape(-1)
    if idx.size != norb.size:
        raise RuntimeError("atom_index must have one entry per atom")
    n_groups = int(idx.max()) + 1 if idx.size else 0
    return PdosGroups(n_groups, np.repeat(idx, norb))


def pdos_groups_by_orbital_slot(atom_type, type_norb) -> PdosGroups:
    """One group per (type t, slot o); group id = offset[t] + o with
    offset[t] = sum of type_norb over the types before t."""
    tn = np.asarray(type_norb, dtype=np.int64).reshape(-1)
    offset = np.concatenate([[0], np.cumsum(tn)])
    n_groups = int(offset[-1])
    group_type = np.repeat(np.arange(tn.size), tn)
    group_slot = np.concatenate([np.arange(n) for n in tn]) if tn.size else np.zeros(0, dtype=np.int64)
    at = np.asarray(atom_type, dtype=np.int64).reshape(-1)
    norb = _norb_per_atom(at, tn)
    rows = [offset[t] + np.arange(n) for t, n in zip(at, norb)]
    gor = np.concatenate(rows) if rows else np.zeros(0, dtype=np.int64)
    return PdosGroups(n_groups, gor, group_type.astype(np.int32), group_slot.astype(np.int32))
