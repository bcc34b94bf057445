"""The b = 1 slab SpMM (a lane owns 16, 32 or 64 contiguous bytes of a row)
against the column-per-lane layout, on uniform rows (graphene: 4 entries; bands
of 9 and 13 entries) and ragged rows,
R = 16 and 32, real and complex vectors, FP64 and FP32: products bit for bit
(the same operations in the same order), moments and adjoint gradients to
1e-12 (only the dots' summation order differs), and FP64 moments against the
oracle."""
import numpy as np
import pytest

from rsatb_amd import _lib
from rsatb_amd.autograd import DifferentiableKPM, sk_pz_parameterization
from rsatb_amd.bsr import BSR
from rsatb_amd.lattice import graphene_rect_supercell, neighbor_edges
from rsatb_amd.models import SkPzModel, assemble

pytestmark = pytest.mark.gpu

SLABS = ("0", "16", "32", "64")  # 0: the column-per-lane layout


@pytest.fixture
def slab_env(monkeypatch):
    def use(v, min_grid="0"):
        monkeypatch.setenv("RG_B1V", v)
        monkeypatch.setenv("RG_B1V_MIN", min_grid)  # 0: slabs at every size
        _lib.Context(0).close()  # the knobs are read at context creation (process-wide)

    yield use
    use("-1", "-1")  # any value but 0, 16, 32, 64 (a negative size): the defaults


def ragged(rng, nbr=700):
    lens = rng.integers(0, 9, size=nbr)
    lens[rng.choice(nbr, 40, replace=False)] = 0
    rows = np.repeat(np.arange(nbr), lens)
    cols = np.concatenate([rng.choice(nbr, k, replace=False) for k in lens])
    vals = rng.standard_normal(rows.size)
    A = BSR.from_blocks(nbr, 1, rows, cols, vals.reshape(-1, 1, 1))
    D = A.to_dense()
    D = 0.5 * (D + D.T)  # symmetric, so the moments are real
    r, c = np.nonzero(D)
    return BSR.from_blocks(nbr, 1, r, c, D[r, c].reshape(-1, 1, 1))


def band(rng, m, nbr=700):
    # symmetric circulant band: every row holds exactly 2 m + 1 entries
    D = np.zeros((nbr, nbr))
    for k in range(1, m + 1):
        v = rng.standard_normal(nbr)
        for i in range(nbr):
            D[i, (i + k) % nbr] = D[(i + k) % nbr, i] = v[i]
    D[np.arange(nbr), np.arange(nbr)] = rng.standard_normal(nbr)
    r, c = np.nonzero(D)
    return BSR.from_blocks(nbr, 1, r, c, D[r, c].reshape(-1, 1, 1))


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("shape", ["graphene", "ragged", "band9", "band13"])
@pytest.mark.parametrize("R", [16, 32])
@pytest.mark.parametrize("samples", ["rademacher", "phase"])
def test_slab_layouts_agree(slab_env, oracle, shape, R, samples):
    rng = np.random.default_rng(R + (3 if shape == "graphene" else 5))
    if shape == "graphene":
        st = graphene_rect_supercell(16, 16)
        H = assemble(st, neighbor_edges(st, 1.6), SkPzModel())
        half = 8.1 / 0.99
    else:
        H = ragged(rng) if shape == "ragged" else band(rng, 4 if shape == "band9" else 6)
        half = 1.01 * np.max(np.abs(np.linalg.eigvalsh(H.to_dense())))
    kind = _lib.SAMPLES_RADEMACHER if samples == "rademacher" else _lib.SAMPLES_PHASE
    X = rng.standard_normal((H.n_rows, R))
    if samples == "phase":
        X = X + 1j * rng.standard_normal((H.n_rows, R))
    out = {}
    for v in SLABS:
        slab_env(v)
        ctx = _lib.Context(0)
        try:
            ctx.upload_bsr(_lib.SLOT_H, H)
            ctx.set_operator(_lib.FORM_ORTHOGONAL, 0.0, half)
            ctx.set_samples(H.n_rows, R, kind, 99)
            y = ctx.bsr_mult(_lib.SLOT_H, X)
            m64 = ctx.kpm_moments(96, per_column=True)
            m32 = ctx.kpm_moments(96, _lib.PREC_FP32, per_column=True)
            mp = ctx.kpm_moments(40, doubling=_lib.DOUBLING_OFF)
        finally:
            ctx.close()
        out[v] = (y, m64, m32, mp)
    ref = out["0"]
    for v in SLABS[1:]:
        y, m64, m32, mp = out[v]
        print(f"\n{shape} R={R} {samples} slab {v}: moments {rel(m64[1], ref[1][1]):.1e} "
              f"fp32 {rel(m32[1], ref[2][1]):.1e} plain {rel(mp, ref[3]):.1e}", flush=True)
        assert np.array_equal(y, ref[0])
        for a, b in ((m64[0], ref[1][0]), (m64[1], ref[1][1]), (m32[0], ref[2][0]), (m32[1], ref[2][1]),
                     (mp, ref[3])):
            assert rel(a, b) <= 1e-12
    # and the default layout against the oracle (FP64 moments)
    draw = oracle.samples_rademacher if samples == "rademacher" else oracle.samples_phase
    mu = oracle.Operator(0, H, 0.0, half).moments(draw(99, H.n_rows, R), 96)
    assert rel(out["16"][1][0], mu) <= 1e-12


def test_slab_adjoint_agrees(slab_env):
    st = graphene_rect_supercell(8, 8)
    model = sk_pz_parameterization(st, neighbor_edges(st, 1.6), [1.6])
    theta = np.array([0.0, -2.7])
    g = np.linspace(1.0, -1.0, 90)
    grads = {}
    for v in SLABS:
        slab_env(v)
        eng = DifferentiableKPM(model, 16, (-8.2, 8.2), 90, checkpoint_stride=8)
        try:
            eng.forward(theta)
            grads[v] = eng.ctx.kpm_moments_backward(g)
        finally:
            eng.ctx.close()
    for v in SLABS[1:]:
        assert rel(grads[v], grads["0"]) <= 1e-12


def test_gradient_accumulation_layouts_agree(monkeypatch):
    # one, two or four sample columns a lane in the pattern contraction
    # dL/dA_ij = sum_c Re(u_i conj v_j): the same sums in another order
    st = graphene_rect_supercell(8, 8)
    model = sk_pz_parameterization(st, neighbor_edges(st, 1.6), [1.6])
    theta = np.array([0.0, -2.7])
    g = np.linspace(1.0, -1.0, 90)
    grads = {}
    try:
        for R in (16, 32):
            for cols in ("1", "2", "4"):
                monkeypatch.setenv("RG_GRAD_WIDE", cols)
                _lib.Context(0).close()  # the knob is read at context creation (process-wide)
                eng = DifferentiableKPM(model, R, (-8.2, 8.2), 90, checkpoint_stride=8)
                try:
                    eng.forward(theta)
                    grads[R, cols] = eng.ctx.kpm_moments_backward(g)
                finally:
                    eng.ctx.close()
    finally:
        monkeypatch.setenv("RG_GRAD_WIDE", "-1")  # the default
        _lib.Context(0).close()
    for R in (16, 32):
        for cols in ("2", "4"):
            assert rel(grads[R, cols], grads[R, "1"]) <= 1e-12
