"""The MFMA SpMM plan (rg_mfma_plan), host only: the row schedule is a
permutation of the block rows in regions of the rows one XCD keeps in
flight, every row's positions take each of its stored blocks once, and the
block order follows one order of the columns shared by all rows, with
columns that share a row apart in it (a distance-2 colouring)."""
import numpy as np
import pytest

from rsatb_amd import _lib
from rsatb_amd.lattice import diamond_supercell, graphene_rect_supercell, neighbor_edges


def pattern(st, cutoff):
    e = neighbor_edges(st, cutoff)
    n = st.n_atoms
    key = np.unique(np.concatenate([e.src * n + e.dst, np.arange(n) * (n + 1)]))
    src, dst = key // n, key % n
    rp = np.zeros(n + 1, np.int64)
    np.add.at(rp, src + 1, 1)
    return np.cumsum(rp), dst.astype(np.int32)


def check_plan(rp, ci):
    n = rp.shape[0] - 1
    sched, kperm, sreg = _lib.mfma_plan(rp, ci)
    assert sreg == max(1, min(128, (n + 7) // 8))
    assert np.array_equal(np.sort(sched), np.arange(n))
    for r in range(n):
        seg = kperm[rp[r]:rp[r + 1]]
        assert np.array_equal(np.sort(seg), np.arange(rp[r], rp[r + 1]))
    return sched, kperm, sreg


@pytest.mark.parametrize("cells", [2, 4])
def test_plan_diamond_second_neighbours(cells):
    rp, ci = pattern(diamond_supercell(cells), 2.7)
    sched, kperm, sreg = check_plan(rp, ci)
    n = rp.shape[0] - 1
    # one order of the columns serves every row: the precedence graph of
    # consecutive positions is acyclic (a topological order exists)
    kc = ci[kperm]
    before = {}
    for r in range(n):
        row = kc[rp[r]:rp[r + 1]]
        for a, b in zip(row[:-1], row[1:]):
            before.setdefault(int(a), set()).add(int(b))
    indeg = np.zeros(n, np.int64)
    for a, bs in before.items():
        for b in bs:
            indeg[b] += 1
    stack = [c for c in range(n) if indeg[c] == 0]
    seen = 0
    while stack:
        a = stack.pop()
        seen += 1
        for b in before.get(a, ()):
            indeg[b] -= 1
            if indeg[b] == 0:
                stack.append(b)
    assert seen == n
    # regions are compact: the neighbour union of a region is far smaller
    # than its rows' block count
    rows = sched[:sreg]
    union = np.unique(np.concatenate([ci[rp[r]:rp[r + 1]] for r in rows]))
    assert union.size < 0.5 * sum(rp[r + 1] - rp[r] for r in rows)


def test_plan_graphene_and_ragged_rows():
    rp, ci = pattern(graphene_rect_supercell(16, 16), 2.5)  # first and second neighbours
    check_plan(rp, ci)
    # ragged rows, an empty row and a disconnected pair
    rp = np.array([0, 3, 3, 5, 6, 7, 8], np.int64)
    ci = np.array([0, 2, 3, 0, 2, 3, 5, 4], np.int32)
    sched, kperm, sreg = check_plan(rp, ci)
    assert sreg == 1


def test_plan_rejects_bad_patterns():
    with pytest.raises(_lib.RsatbGpuError):
        _lib.mfma_plan(np.array([0, 1], np.int64), np.array([1], np.int32))  # column out of range
    with pytest.raises(_lib.RsatbGpuError):
        _lib.mfma_plan(np.array([0, 2, 1], np.int64), np.array([0, 1], np.int32))  # decreasing row_ptr
