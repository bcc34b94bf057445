#!/usr/bin/env python3
"""Benchmark of the stochastic Chebyshev/KPM moment loop on one MI355X.

Workload (one "step" = one full moment computation, kpm_dos_moments of
rsatb/calc/dos.hpp:124-189, over all sample columns):
  diamond, 8 x 8 x 8 conventional cells (4096 atoms), 13 orbitals per atom,
  random symmetric H and S blocks to the 2.7 A second-neighbour cutoff, the
  sandwich operator Z^H H Z with Z = graph-local S^-1/2 built on the GPU,
  spectral window from GPU Lanczos + 0.5 eV padding, R = 64 random-phase
  complex vectors, M = 2048 moments, FP64.
The value is orbital x vector x moment units per second with all operands
resident in HBM. Per-kernel times come from HIP events recorded on the
library's stream over the timed steps; the CPU baseline is the oracle
(the FP64 C restatement of the reference's loop) on a bounded sample.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision fp64|fp32]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "Chebyshev orbital·vector·moment steps/s, FP32-MFMA vs FP64 at ≤1e-5 DOS err; bwd/fwd"
UNIT = "orbital*vector*moment/s"
HBM_PEAK_GBS = 8000.0
PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--precision", choices=["fp64", "fp32"], default="fp64")
    p.add_argument("--cells", type=int, default=8, help="diamond conventional cells per edge")
    p.add_argument("--nsamples", type=int, default=64)
    p.add_argument("--moments", type=int, default=2048)
    p.add_argument("--cpu-moments", type=int, default=32, help="moments in the timed CPU sample")
    p.add_argument("--no-cpu-baseline", action="store_true")
    return p.parse_args()


def roofline(stats: dict, precision: str) -> dict:
    """The dominant kernel class (largest total time) against its roofline."""
    best = max((s for s in stats.values() if s["count"] > 0), key=lambda s: s["avg_ms"] * s["count"])
    t = best["avg_ms"] * 1e-3
    intensity = best["flops"] / best["bytes"]
    ridge = PEAK_TFLOPS[precision] * 1e12 / (HBM_PEAK_GBS * 1e9)
    gbs = best["bytes"] / t / 1e9
    tfs = best["flops"] / t / 1e12
    if intensity > ridge:
        out = {"bound": "mfma", "achieved": round(tfs, 3), "peak": PEAK_TFLOPS[precision], "unit": "TFLOP/s",
               "frac": round(tfs / PEAK_TFLOPS[precision], 4)}
    else:
        out = {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
               "frac": round(gbs / HBM_PEAK_GBS, 4)}
    out.update({"traffic": None, "kernel": best["name"], "avg_ms": round(best["avg_ms"], 5),
                "algorithmic_bytes": best["bytes"], "algorithmic_flops": best["flops"],
                "intensity_flop_per_byte": round(intensity, 3), "achieved_GBs": round(gbs, 1),
                "achieved_TFLOPs": round(tfs, 3)})
    return out


def cpu_baseline(calc, n_moments: int, ref_mu) -> dict:
    """The oracle's fused FP64 loop on the host cores over the same operator,
    the same samples and the first n_moments moments."""
    import numpy as np

    from oracle import oracle as O
    from rsatb_amd import _lib

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    O.set_threads(threads)
    ctx = calc.context
    ks = calc._kstates[0]
    emin, emax = calc.energy_bounds()
    c, a = (emax + emin) / 2, (emax - emin) / 2
    if ks["S"] is not None:
        Z = ctx.download_bsr(_lib.SLOT_Z)
        op = O.Operator(O.FORM_SANDWICH, ks["H"], c, a, Z, Z.conj_transpose())
    else:
        op = O.Operator(O.FORM_ORTHOGONAL, ks["H"], c, a)
    psi = ctx.get_samples()
    op.moments(psi, 4, doubling=True, fused=True)  # warm-up
    t0 = time.perf_counter()
    mu = op.moments(psi, n_moments, doubling=True, fused=True)
    dt = time.perf_counter() - t0
    rows, R = psi.shape
    err = float(np.max(np.abs(mu - ref_mu[:n_moments])))
    return {"value": rows * R * n_moments / dt, "unit": UNIT, "cores": threads, "kind": "port",
            "sample": f"same operator and samples, first {n_moments} moments ({n_moments // 2} sandwich applies, "
                      f"doubled), fused FP64 oracle loop, {dt:.2f} s",
            "seconds": round(dt, 3), "max_abs_moment_diff_vs_gpu": err}


def main() -> int:
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported (the moment loop shards "
              "sample columns across GPUs by design, but no multi-GPU run is part of this build)", file=sys.stderr)
        return 2
    import numpy as np

    from rsatb_amd import _lib
    from rsatb_amd.driver import RSATB, PreconditionerOptions, SetupOptions
    from rsatb_amd.lattice import diamond_supercell
    from rsatb_amd.models import RandomBlockModel

    t_setup = time.perf_counter()
    st = diamond_supercell(args.cells)
    calc = RSATB(st, RandomBlockModel(), SetupOptions(
        nsamples=args.nsamples, sample_seed=20240917, compute_bounds=True, sample_kind="phase",
        preconditioner=PreconditionerOptions(operator_form="sandwich"), precision=args.precision))
    calc._bind_kpoint(0)
    ctx = calc.context
    setup_s = time.perf_counter() - t_setup
    prec = _lib.PREC_FP32 if args.precision == "fp32" else _lib.PREC_FP64
    rows = st.n_atoms * 13
    M = args.moments
    print(f"setup {setup_s:.1f}s: {rows} orbitals, bounds {calc.energy_bounds()}, patches {calc._kstates[0].get('isqrt')}",
          file=sys.stderr, flush=True)

    for _ in range(args.warmup):
        mu = ctx.kpm_moments(M, prec)
    ctx.set_profiling(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        mu = ctx.kpm_moments(M, prec)  # synchronous on return
    dt = time.perf_counter() - t0
    stats = {}
    for k, name in _lib.KCLASS_NAMES.items():
        s = ctx.kernel_stats(k)
        s["name"] = name
        stats[name] = s
    ctx.set_profiling(False)
    ms = dt / args.steps * 1e3
    value = rows * args.nsamples * M / (dt / args.steps)

    line = {
        "metric": METRIC, "value": value, "unit": UNIT, "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32 (f64 dots)", "data": "synthetic",
        "config": {"workload": f"diamond {args.cells}x{args.cells}x{args.cells} conventional cells "
                               f"({st.n_atoms} atoms x 13 orbitals), sandwich Z^H H Z, R={args.nsamples} "
                               f"random-phase complex vectors, M={M} moments",
                   "n_orbitals": rows, "R": args.nsamples, "M": M, "form": "sandwich",
                   "precision": args.precision, "bounds": list(calc.energy_bounds())},
        "roofline": roofline(stats, args.precision),
        "kernels": {k: {"avg_ms": round(v["avg_ms"], 5), "count": v["count"]} for k, v in stats.items()},
        "mu0": float(mu[0].real), "max_abs_mu": float(np.max(np.abs(mu))),
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(calc, args.cpu_moments, mu)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
