"""The C ABI library: loads, exports every declared symbol, host-only entry
points agree with the oracle (no device calls here)."""
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    text = (ROOT / "include" / "rsatb_gpu.h").read_text()
    return sorted(set(re.findall(r"RG_API\s+[\w\s\*]*?\b(rg_\w+)\s*\(", text)))


def test_header_declares_expected_entry_points():
    syms = declared_symbols()
    for name in ["rg_create", "rg_destroy", "rg_upload_bsr", "rg_set_operator", "rg_set_samples",
                 "rg_apply_op", "rg_kpm_moments", "rg_compute_dos", "rg_cancel", "rg_last_error"]:
        assert name in syms


def test_library_exports_every_declared_symbol():
    from rsatb_amd import _lib

    lib = _lib.load()
    for name in declared_symbols():
        assert hasattr(lib, name), name
        assert name in _lib.SIGNATURES, f"{name} has no ctypes signature"
    assert lib.rg_version() >= 1


def test_dos_plan_matches_oracle(oracle):
    from rsatb_amd import _lib

    for emin, emax, de in [(-8.6, 8.6, 0.15), (-3.0, 11.0, 0.01), (0.0, 1.0, 2.0)]:
        npts, n = _lib.dos_plan(emin, emax, de)
        nt, n2 = oracle.dos_plan(emin, emax, de)
        assert npts == 2 * nt and n == n2
    with pytest.raises(_lib.RsatbGpuError):
        _lib.dos_plan(1.0, 1.0, 0.1)
    with pytest.raises(_lib.RsatbGpuError):
        _lib.dos_plan(-1.0, 1.0, 0.0)


@pytest.mark.parametrize("kernel", [0, 1, 2, 3])
def test_kpm_density_matches_oracle(oracle, kernel):
    from rsatb_amd import _lib

    rng = np.random.default_rng(kernel)
    mu = rng.standard_normal(77) + 1j * rng.standard_normal(77)
    x = np.linspace(-1.0, 1.0, 301)
    a = _lib.kpm_density(mu, x, 0.5 / 150, kernel)
    b = oracle.kpm_density(mu, x, 0.5 / 150, kernel)
    assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(b))


def test_no_device_gives_clear_error():
    if Path("/dev/kfd").exists():
        pytest.skip("a GPU driver is present")
    from rsatb_amd import _lib

    with pytest.raises(_lib.RsatbGpuError):
        _lib.Context(0)
