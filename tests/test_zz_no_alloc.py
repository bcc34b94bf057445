"""rg_kpm_moments_noalloc (the moment loop with its workspace in the
library's fixed device array, no device allocation) against rg_kpm_moments,
through the C ABI: the same kernels in the same order, so the moments agree
bit for bit. Two calls in a row share the array, the second on a smaller
input; the orthogonal, sandwich, plain-recurrence and FP32 forms; a workspace
beyond the array is rejected before any launch."""
import numpy as np
import pytest

from rsatb_amd import _lib
from rsatb_amd.bsr import BSR
from rsatb_amd.lattice import diamond_supercell, graphene_rect_supercell, neighbor_edges
from rsatb_amd.models import RandomBlockModel, SkPzModel, assemble

pytestmark = pytest.mark.gpu


def graphene(cells):
    st = graphene_rect_supercell(cells, cells)
    return assemble(st, neighbor_edges(st, 1.6), SkPzModel())


def both(ctx, M, prec=_lib.PREC_FP64, doubling=_lib.DOUBLING_AUTO):
    ref, ref_cols = ctx.kpm_moments(M, prec, doubling, per_column=True)
    got, got_cols = ctx.kpm_moments_noalloc(M, prec, doubling, per_column=True)
    return ref, ref_cols, got, got_cols


def test_noalloc_matches_allocating_call_twice_shrinking():
    ctx = _lib.Context(0)
    try:
        for cells, R, M in ((16, 16, 256), (8, 8, 64)):
            H = graphene(cells)
            ctx.upload_bsr(_lib.SLOT_H, H)
            ctx.set_operator(_lib.FORM_ORTHOGONAL, 0.0, 8.1 / 0.99)
            ctx.set_samples(H.n_rows, R, _lib.SAMPLES_RADEMACHER, 20240917)
            print(f"\ngraphene {cells}x{cells}, R={R}, M={M}", flush=True)
            ref, rc_, got, gc_ = both(ctx, M)
            assert np.array_equal(ref, got) and np.array_equal(rc_, gc_)
            # the plain recurrence (psi copy in the array too)
            ref, rc_, got, gc_ = both(ctx, M // 2, doubling=_lib.DOUBLING_OFF)
            assert np.array_equal(ref, got) and np.array_equal(rc_, gc_)
    finally:
        ctx.close()


def test_noalloc_sandwich_and_fp32():
    st = diamond_supercell(2)
    e = neighbor_edges(st, 2.7)
    m = RandomBlockModel()
    ctx = _lib.Context(0)
    try:
        ctx.upload_bsr(_lib.SLOT_S, assemble(st, e, m, "S"))
        ctx.isqrt_patches()
        H = assemble(st, e, m, "H")
        ctx.upload_bsr(_lib.SLOT_H, H)
        lo, hi = ctx.estimate_bounds(_lib.FORM_SANDWICH, steps=100, padding=0.5)
        ctx.set_operator(_lib.FORM_SANDWICH, (hi + lo) / 2, (hi - lo) / 2)
        for R, M in ((32, 128), (8, 40)):
            ctx.set_samples(H.n_rows, R, _lib.SAMPLES_PHASE, 7)
            print(f"\ndiamond 2^3 sandwich, R={R}, M={M}", flush=True)
            ref, rc_, got, gc_ = both(ctx, M)
            assert np.array_equal(ref, got) and np.array_equal(rc_, gc_)
            ref, rc_, got, gc_ = both(ctx, M, _lib.PREC_FP32)  # FP32 copies made by the first call
            assert np.array_equal(ref, got) and np.array_equal(rc_, gc_)
        # new values into the same slot: the no-allocation call runs first and
        # remakes the value copies made earlier (FP32 values, the MFMA products'
        # ordered values) in their buffers; the allocating call after it reads
        # the same copies
        H2 = assemble(st, e, RandomBlockModel(seed=3), "H")
        ctx.upload_bsr(_lib.SLOT_H, H2)
        for prec in (_lib.PREC_FP64, _lib.PREC_FP32):
            got, gc_ = ctx.kpm_moments_noalloc(40, prec, per_column=True)
            ref, rc_ = ctx.kpm_moments(40, prec, per_column=True)
            print(f"\nre-uploaded H, precision {prec}: max |diff| {np.max(np.abs(got - ref)):.1e}", flush=True)
            assert np.array_equal(ref, got) and np.array_equal(rc_, gc_)
    finally:
        ctx.close()


def test_noalloc_rejects_oversized_workspace_and_missing_fp32_copies():
    ctx = _lib.Context(0)
    try:
        H = graphene(256)  # 262,144 rows: 3 real vectors of R = 64 exceed 128 MiB
        ctx.upload_bsr(_lib.SLOT_H, H)
        ctx.set_operator(_lib.FORM_ORTHOGONAL, 0.0, 8.1 / 0.99)
        ctx.set_samples(H.n_rows, 64, _lib.SAMPLES_RADEMACHER, 1)
        with pytest.raises(_lib.RsatbGpuError, match="scratch") as exc:
            ctx.kpm_moments_noalloc(16)
        assert exc.value.code == _lib.RG_ERR_UNSUPPORTED
        H = graphene(8)
        ctx.upload_bsr(_lib.SLOT_H, H)  # a new H: its FP32 copy does not exist yet
        ctx.set_samples(H.n_rows, 4, _lib.SAMPLES_RADEMACHER, 1)
        with pytest.raises(_lib.RsatbGpuError, match="FP32") as exc:
            ctx.kpm_moments_noalloc(16, _lib.PREC_FP32)
        assert exc.value.code == _lib.RG_ERR_STATE
        # and the array still serves a valid call afterwards
        ref = ctx.kpm_moments(32)
        assert np.array_equal(ref, ctx.kpm_moments_noalloc(32))
    finally:
        ctx.close()
