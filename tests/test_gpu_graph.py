"""Graph replay of the moment loop (small operators): the captured and
replayed launch sequence gives the eager loop's moments bit for bit, on every
replay, with eager work and synchronisations in between; a re-upload into the
same buffers replays with the new values; a changed configuration
recaptures."""
import numpy as np
import pytest

from rsatb_amd import _lib
from rsatb_amd.lattice import graphene_rect_supercell, neighbor_edges
from rsatb_amd.models import SkPzModel, assemble

pytestmark = pytest.mark.gpu


def _ctx(monkeypatch, mode):
    monkeypatch.setenv("RG_GRAPH", str(mode))
    return _lib.Context(0)


def test_graph_replays_match_eager(monkeypatch):
    st = graphene_rect_supercell(32, 32)
    H = assemble(st, neighbor_edges(st, 1.6), SkPzModel())
    N, R, M = st.n_atoms, 16, 512
    eager, graph = _ctx(monkeypatch, 0), _ctx(monkeypatch, 1)
    try:
        for c in (eager, graph):
            c.upload_bsr(_lib.SLOT_H, H)
            c.set_operator(_lib.FORM_ORTHOGONAL, 0.0, 8.1 / 0.99)
            c.set_samples(N, R, _lib.SAMPLES_RADEMACHER, 20240917)
        ref, ref_cols = eager.kpm_moments(M, per_column=True)
        for rep in range(4):  # capture, then replays
            mu, cols = graph.kpm_moments(M, per_column=True)
            print(f"\nreplay {rep}: max |diff| {np.max(np.abs(cols - ref_cols)):.1e}", flush=True)
            assert np.array_equal(cols, ref_cols) and np.array_equal(mu, ref)
            eager.kpm_moments(64)  # eager work between the replays
        # new values in the same buffers: the replay uses them
        H2 = assemble(st, neighbor_edges(st, 1.6), SkPzModel(shells=[(1.6, -2.6, 0.0)]))
        for c in (eager, graph):
            c.upload_bsr(_lib.SLOT_H, H2)
        ref2 = eager.kpm_moments(M)
        assert np.array_equal(graph.kpm_moments(M), ref2)
        assert not np.array_equal(ref2, ref)
        # a new window and a new M: recaptured
        for c in (eager, graph):
            c.set_operator(_lib.FORM_ORTHOGONAL, 0.1, 8.5)
        assert np.array_equal(graph.kpm_moments(M + 130), eager.kpm_moments(M + 130))
        # FP32 form and the checkpointed forward of the adjoint
        assert np.array_equal(graph.kpm_moments(M, _lib.PREC_FP32), eager.kpm_moments(M, _lib.PREC_FP32))
        for c in (eager, graph):
            c.set_checkpointing(8)
        g = np.linspace(1.0, -1.0, 96)
        grads = []
        for c in (eager, graph):
            c.kpm_moments(96)
            grads.append(c.kpm_moments_backward(g))
        assert np.array_equal(grads[0], grads[1])
    finally:
        eager.close()
        graph.close()
