"""The precision-form headline configuration at its full size: diamond
24 x 24 x 24 conventional cells (110,592 atoms x 13 orbitals = 1,437,696
orbitals), sandwich Z^H H Z with the graph-local Z = S^-1/2 built on the GPU,
R = 128 random-phase complex vectors. The first moments are checked against
the FP64 oracle on the host cores (two sandwich applies of the same operator,
Z downloaded from the device, the same samples); the rest through properties
that hold at any size: mu_0 = 1 for unit-norm samples, |mu_n| <= 1 inside the
window, and the FP32 form within its error budget of FP64. This runs the
MFMA kernels with their row schedule and colour-ordered blocks on the full
pattern."""
import os

import numpy as np
import pytest

from rsatb_amd import _lib

pytestmark = pytest.mark.gpu


def test_c3_full_size_against_oracle_and_properties(oracle):
    from rsatb_amd.driver import RSATB, PreconditionerOptions, SetupOptions
    from rsatb_amd.lattice import diamond_supercell
    from rsatb_amd.models import RandomBlockModel

    st = diamond_supercell(24)
    calc = RSATB(st, RandomBlockModel(), SetupOptions(
        nsamples=128, sample_seed=20240917, compute_bounds=True, sample_kind="phase",
        preconditioner=PreconditionerOptions(operator_form="sandwich")))
    calc._bind_kpoint(0)
    ctx = calc.context
    print(f"\ndiamond 24^3 sandwich, {st.n_atoms * 13} orbitals, bounds {calc.energy_bounds()}", flush=True)
    M = 64
    mu64 = ctx.kpm_moments(M, _lib.PREC_FP64)
    mu32 = ctx.kpm_moments(M, _lib.PREC_FP32)
    assert abs(mu64[0] - 1.0) < 1e-12
    assert np.all(np.abs(mu64) <= 1.0 + 1e-12)
    assert np.max(np.abs(mu32 - mu64)) < 1e-7
    # the oracle: first 4 moments (two doubled sandwich applies) on the host
    oracle.set_threads(int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))
    ks = calc._kstates[0]
    emin, emax = calc.energy_bounds()
    Z = ctx.download_bsr(_lib.SLOT_Z)
    op = oracle.Operator(oracle.FORM_SANDWICH, ks["H"], (emax + emin) / 2, (emax - emin) / 2, Z,
                         Z.conj_transpose())
    ref = op.moments(ctx.get_samples(), 4, doubling=True, fused=True)
    d = float(np.max(np.abs(mu64[:4] - ref)))
    print(f"max |mu_gpu - mu_oracle| over 4 moments: {d:.2e}", flush=True)
    assert d < 1e-10
    ctx.close()
