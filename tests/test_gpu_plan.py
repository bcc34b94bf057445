"""The MFMA SpMM's plan on the device (row schedule and colour-ordered
blocks, on by default for real slots with b >= 8) against the stored order
(RG_MFMA_PLAN=0) and the oracle: the same products and moments to FP64
rounding; values written after a run (a re-upload, new patches) reach the
next run, also when it replays a captured graph."""
import numpy as np
import pytest

from rsatb_amd import _lib
from rsatb_amd.bsr import BSR
from rsatb_amd.lattice import diamond_supercell, neighbor_edges
from rsatb_amd.models import RandomBlockModel, assemble

pytestmark = pytest.mark.gpu


def _ctx(monkeypatch, plan, graph=2):
    monkeypatch.setenv("RG_MFMA_PLAN", "1" if plan else "0")
    monkeypatch.setenv("RG_GRAPH", str(graph))
    return _lib.Context(0)


def _diamond(cells, seed=0):
    st = diamond_supercell(cells)
    e = neighbor_edges(st, 2.7)
    m = RandomBlockModel(seed=seed) if seed else RandomBlockModel()
    return st, assemble(st, e, m, "H"), assemble(st, e, m, "S")


def _setup(ctx, H, S, R, seed=20240917):
    ctx.upload_bsr(_lib.SLOT_S, S)
    ctx.isqrt_patches()
    ctx.upload_bsr(_lib.SLOT_H, H)
    lo, hi = ctx.estimate_bounds(_lib.FORM_SANDWICH, steps=60, padding=0.5)
    ctx.set_operator(_lib.FORM_SANDWICH, (hi + lo) / 2, (hi - lo) / 2)
    ctx.set_samples(H.n_rows, R, _lib.SAMPLES_PHASE, seed)
    return lo, hi


def test_plan_products_match_stored_order_and_oracle(monkeypatch, oracle):
    st, H, _ = _diamond(4)  # 512 atoms: 2 regions of 32 rows per XCD
    rng = np.random.default_rng(3)
    X = rng.standard_normal((H.n_rows, 64)) + 1j * rng.standard_normal((H.n_rows, 64))
    ref = oracle.bsr_mult(H, X)
    ys = []
    for plan in (True, False):
        ctx = _ctx(monkeypatch, plan)
        try:
            ctx.upload_bsr(_lib.SLOT_H, H)
            print(f"\ndiamond 4^3 H x, plan={plan}", flush=True)
            ys.append(ctx.bsr_mult(_lib.SLOT_H, X))
        finally:
            ctx.close()
    for y in ys:
        assert np.max(np.abs(y - ref)) <= 1e-12 * np.max(np.abs(ref))


@pytest.mark.parametrize("R", [64, 128])
def test_plan_moments_match_stored_order(monkeypatch, R):
    st, H, S = _diamond(4)
    mus, win = {}, []
    for plan in (True, False):
        ctx = _ctx(monkeypatch, plan, graph=0)
        try:
            lo, hi = _setup(ctx, H, S, R)
            print(f"\ndiamond 4^3 sandwich R={R} plan={plan} bounds ({lo:.6f}, {hi:.6f})", flush=True)
            if not win:  # one window for both, from the first estimate
                win = [(hi + lo) / 2, (hi - lo) / 2 + 0.5]
            ctx.set_operator(_lib.FORM_SANDWICH, *win)
            mus[plan] = (ctx.kpm_moments(256), ctx.kpm_moments(256, _lib.PREC_FP32))
        finally:
            ctx.close()
    assert np.max(np.abs(mus[True][0] - mus[False][0])) < 1e-13
    assert np.max(np.abs(mus[True][1] - mus[False][1])) < 1e-6
    assert np.max(np.abs(mus[True][1] - mus[True][0])) < 1e-5


def test_plan_values_follow_writes_and_graph_replays(monkeypatch):
    # a replayed graph reads the ordered value copies: a re-upload of H and
    # new patches from a new S must reach them before the replay
    st, H, S = _diamond(2)
    _, H2, S2 = _diamond(2, seed=11)
    ctxs = [_ctx(monkeypatch, True, graph=1), _ctx(monkeypatch, False, graph=0)]
    try:
        lo, hi = _setup(ctxs[0], H, S, 32)
        _setup(ctxs[1], H, S, 32)
        for c in ctxs:  # wide enough for both operators
            c.set_operator(_lib.FORM_SANDWICH, (hi + lo) / 2, (hi - lo) / 2 + 4.0)
        M = 200
        a = [c.kpm_moments(M) for c in ctxs]
        b = [c.kpm_moments(M) for c in ctxs]  # the planned context replays
        print(f"\nfirst/replay: {np.max(np.abs(a[0] - a[1])):.1e} {np.max(np.abs(b[0] - b[1])):.1e}", flush=True)
        assert np.array_equal(a[0], b[0]) and np.max(np.abs(a[0] - a[1])) < 1e-13
        for c in ctxs:
            c.upload_bsr(_lib.SLOT_H, H2)
        a2 = [c.kpm_moments(M) for c in ctxs]
        print(f"new H: {np.max(np.abs(a2[0] - a2[1])):.1e}", flush=True)
        assert np.max(np.abs(a2[0] - a2[1])) < 1e-13 and np.max(np.abs(a2[0] - a[0])) > 1e-6
        for c in ctxs:
            c.upload_bsr(_lib.SLOT_S, S2)
            c.isqrt_patches()
        a3 = [c.kpm_moments(M) for c in ctxs]
        a3_32 = [c.kpm_moments(M, _lib.PREC_FP32) for c in ctxs]
        print(f"new S: {np.max(np.abs(a3[0] - a3[1])):.1e}", flush=True)
        assert np.max(np.abs(a3[0] - a3[1])) < 1e-13 and np.max(np.abs(a3[0] - a2[0])) > 1e-6
        assert np.max(np.abs(a3_32[0] - a3_32[1])) < 1e-6
    finally:
        for c in ctxs:
            c.close()
