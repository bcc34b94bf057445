#!/usr/bin/env python3
"""Benchmark of the stochastic Chebyshev/KPM moment loop on one MI355X.

Headline workload (the largest single-GPU configuration):
  diamond, 24 x 24 x 24 conventional cells (110,592 atoms, 1,437,696
  orbitals), 13 orbitals per atom, random symmetric H and S blocks to the
  2.7 A second-neighbour cutoff, the sandwich operator Z^H H Z with Z the
  graph-local S^-1/2 built on the GPU, spectral window from GPU Lanczos +
  0.5 eV padding, R = 128 random-phase complex vectors, M = 4096 moments.
One timed step is a fixed slice of that recurrence at full N and R: the
first 128 moments (64 doubled sandwich applies, kpm_dos_moments of
rsatb/calc/dos.hpp:124-189 over all sample columns), FP64. The complete
M = 4096 runs in FP64 and in the FP32 form (FP32 MFMA products and
recurrence, FP64 dots) run once outside the timed steps and give the
per-moment and DOS errors of the FP32 form. The value is orbital x vector x
moment units per second with all operands resident in HBM. Per-kernel times
come from HIP events recorded on the library's stream in a second pass over
the same steps; the CPU baseline is the oracle (the FP64 C restatement of
the reference's loop, in the reference's unfused form over complex storage
and fused) on a bounded sample of the same workload. Secondary keys: the
diamond 8^3 configuration (R = 64, M = 2048), the graphene 64 x 64 plain DOS
(R = 16, M = 1024), the differentiable graphene 128 x 128 form with a
learnable overlap, SK assembly, Kubo-Greenwood, the b = 1 graphene scale-up
and the twisted-bilayer transport leg.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
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
    p.add_argument("--cells", type=int, default=24, help="diamond conventional cells per edge (headline)")
    p.add_argument("--nsamples", type=int, default=128)
    p.add_argument("--step-moments", type=int, default=128, help="moments per timed step (a slice of the run)")
    p.add_argument("--moments", type=int, default=4096, help="moments of the complete FP64 / FP32 runs")
    p.add_argument("--no-full", action="store_true", help="skip the complete FP64 / FP32 runs")
    p.add_argument("--cpu-moments", type=int, default=4,
                   help="moments in each timed CPU sample (2 sandwich applies at the headline size)")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--no-fp32", action="store_true", help="skip the FP32 steps")
    p.add_argument("--fp32-only", action="store_true",
                   help="time the steps in the FP32 form only (profiling runs; the line's dtype says f32)")
    p.add_argument("--no-legs", action="store_true", help="only the headline")
    p.add_argument("--leg", default=None,
                   choices=["c2", "c1", "adjoint", "sk_assembly", "kubo_greenwood", "graphene_scale", "tblg_transport"],
                   help="run only this secondary leg (no headline) and print its record")
    p.add_argument("--no-c2", action="store_true", help="skip the diamond 8^3 leg")
    p.add_argument("--no-c1", action="store_true", help="skip the graphene 64x64 plain-DOS leg")
    p.add_argument("--no-adjoint", action="store_true", help="skip the backward/forward leg")
    p.add_argument("--adj-cells", type=int, default=128, help="graphene rectangular cells per edge (adjoint leg)")
    p.add_argument("--no-sk", action="store_true", help="skip the Slater-Koster assembly leg")
    p.add_argument("--sk-cells", type=int, default=16, help="diamond Si cells per edge (SK assembly leg)")
    p.add_argument("--sk-cpu-edges", type=int, default=20000, help="edge blocks the CPU oracle evaluates")
    p.add_argument("--no-kg", action="store_true", help="skip the Kubo-Greenwood leg")
    p.add_argument("--kg-cells", type=int, default=128, help="graphene rectangular cells per edge (KG leg)")
    p.add_argument("--kg-samples", type=int, default=16)
    p.add_argument("--kg-moments", type=int, default=512)
    p.add_argument("--kg-cpu-moments", type=int, default=128)
    p.add_argument("--no-graphene-scale", action="store_true", help="skip the b=1 graphene scale-up leg")
    p.add_argument("--graphene-cells", type=int, default=512, help="rectangular cells per edge (b=1 scale-up leg)")
    p.add_argument("--no-tblg", action="store_true", help="skip the differentiable twisted-bilayer transport leg")
    p.add_argument("--tblg-reps", type=int, default=4, help="1.05-degree moire cells per edge (tBLG leg)")
    p.add_argument("--tblg-samples", type=int, default=16)
    p.add_argument("--tblg-moments", type=int, default=512)
    return p.parse_args()


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


PROFILES = Path(__file__).resolve().parent / "profiles"


def measured_traffic(kclass: str, precision: str, workload_tag: str):
    """HBM bytes per launch of a kernel class, from the committed rocprofv3
    counter summary of this workload and precision (FETCH_SIZE + WRITE_SIZE,
    calibrated on a copy kernel), or None when no summary covers it."""
    try:
        rec = json.loads((PROFILES / f"kpm_{workload_tag}_{precision}.json").read_text())
    except (OSError, ValueError):
        return None
    if rec.get("workload_tag") != workload_tag:
        return None
    dtype = "f32" if precision == "fp32" else "f64"
    best = None
    for k in rec.get("kernels", {}).values():
        if k.get("class") == kclass and k.get("dtype") == dtype and "traffic_bytes" in k:
            if best is None or k["calls"] > best["calls"]:
                best = k
    return None if best is None else round(best["traffic_bytes"])


def roofline(stats: dict, precision: str, workload_tag: str = "") -> dict:
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
    out.update({"traffic": measured_traffic(best["name"], precision, workload_tag), "kernel": best["name"],
                "avg_ms": round(best["avg_ms"], 5),
                "algorithmic_bytes": best["bytes"], "algorithmic_flops": best["flops"],
                "intensity_flop_per_byte": round(intensity, 3), "achieved_GBs": round(gbs, 1),
                "achieved_TFLOPs": round(tfs, 3)})
    return out


def cpu_baseline(calc, n_moments: int, ref_mu) -> dict:
    """The oracle's FP64 loop on the host cores over the same operator and
    the same samples, first n_moments moments, in the two modes the
    reference's CPU path is measured in: the reference's own form (value;
    unfused passes over complex<double> storage, rsatb/calc/dos.hpp:169-183,
    cheby.hpp:589-590) and the fused loop."""
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
    rows, R = psi.shape
    out = {}
    for name, fused in (("reference_form", False), ("fused", True)):
        t0 = time.perf_counter()
        mu = op.moments(psi, n_moments, doubling=True, fused=fused)
        dt = time.perf_counter() - t0
        out[name] = {"value": rows * R * n_moments / dt, "seconds": round(dt, 3),
                     "max_abs_moment_diff_vs_gpu": float(np.max(np.abs(mu - ref_mu[:n_moments])))}
        log(f"cpu baseline {name}: {dt:.2f} s")
    applies = 1 + max(0, (n_moments - 2) // 2)
    return {"value": out["reference_form"]["value"], "unit": UNIT, "cores": threads, "kind": "port",
            "sample": f"same operator and samples, first {n_moments} moments ({applies} sandwich applies, doubled), "
                      f"FP64 oracle on {threads} host threads, no warm-up; value = the reference's unfused form over "
                      f"complex<double> storage ({out['reference_form']['seconds']} s), fused loop beside it "
                      f"({out['fused']['seconds']} s)",
            "reference_form": out["reference_form"], "fused": out["fused"]}


def dos_error(mu_ref, mu):
    """max_E |rho - rho_ref| / max_E rho_ref on 2001 points of [-1, 1]."""
    import numpy as np

    from rsatb_amd import _lib

    x = -1.0 + np.arange(2001) / 1000.0
    r_ref = _lib.kpm_density(mu_ref, x, 0.5 / 1000.0)
    r = _lib.kpm_density(mu, x, 0.5 / 1000.0)
    return float(np.max(np.abs(r - r_ref)) / np.max(r_ref))


def adjoint_leg(cells: int, steps: int) -> dict:
    """Backward/forward time of the differentiable moments on graphene
    (cells x cells rectangular, onsite + 3 hopping shells and the
    nearest-neighbour overlap s1 = 0.1, all learnable), R = 32, M = 512,
    checkpoints every 32 steps. A forward builds S(s1), its graph-local factor
    Z on the device and the sandwich moments; a backward runs the adjoint
    recurrence, dL/dH, dL/dZ and dL/dS through the patches."""
    import numpy as np

    from rsatb_amd import _lib
    from rsatb_amd.autograd import DifferentiableKPM, sk_pz_overlap_parameterization, sk_pz_parameterization
    from rsatb_amd.lattice import graphene_rect_supercell, neighbor_edges

    st = graphene_rect_supercell(cells, cells)
    model = sk_pz_parameterization(st, neighbor_edges(st, 2.9), [1.6, 2.5, 2.9])
    overlap = sk_pz_overlap_parameterization(st, neighbor_edges(st, 1.6), [1.6])
    theta = np.array([0.0, -2.7, -0.2, -0.18, 0.1])
    with _lib.Context(0) as c:
        c.upload_bsr(_lib.SLOT_S, overlap.build(theta[4:]))
        c.isqrt_patches()
        c.upload_bsr(_lib.SLOT_H, model.build(theta[:4]))
        lo, hi = c.estimate_bounds(_lib.FORM_SANDWICH, steps=200, padding=0.5)
    M = 512
    eng = DifferentiableKPM(model, 32, (lo, hi), M, checkpoint_stride=32, sample_kind="phase", overlap=overlap)
    g = np.linspace(1.0, -1.0, M)
    eng.forward(theta)
    eng.backward(g)
    t0 = time.perf_counter()
    for _ in range(steps):
        eng.forward(theta)
    t_f = (time.perf_counter() - t0) / steps
    t_b = 0.0
    for _ in range(steps):
        eng.forward(theta)
        t1 = time.perf_counter()
        grad = eng.backward(g)
        t_b += time.perf_counter() - t1
    t_b /= steps
    # the same pair with the factor held fixed (no overlap gradient)
    Zfix = None
    with _lib.Context(0) as c:
        c.upload_bsr(_lib.SLOT_S, overlap.build(theta[4:]))
        c.isqrt_patches()
        Zfix = c.download_bsr(_lib.SLOT_Z)
    eng.ctx.close()
    eng2 = DifferentiableKPM(model, 32, (lo, hi), M, checkpoint_stride=32, sample_kind="phase", Z=Zfix)
    eng2.forward(theta[:4])
    eng2.backward(g)
    t0 = time.perf_counter()
    for _ in range(steps):
        eng2.forward(theta[:4])
    t_f2 = (time.perf_counter() - t0) / steps
    t_b2 = 0.0
    for _ in range(steps):
        eng2.forward(theta[:4])
        t1 = time.perf_counter()
        eng2.backward(g)
        t_b2 += time.perf_counter() - t1
    t_b2 /= steps
    eng2.ctx.close()
    return {"workload": f"graphene {cells}x{cells} rectangular ({st.n_atoms} atoms), onsite + 3 hopping shells + "
                        "NN overlap s1=0.1 (all learnable; S^-1/2 patches rebuilt in every forward), R=32 phase, "
                        "M=512, checkpoints every 32 steps, FP64",
            "forward_ms": round(t_f * 1e3, 3), "backward_ms": round(t_b * 1e3, 3),
            "bwd_fwd": round(t_b / t_f, 3), "grad_theta": [float(x) for x in grad],
            "params": list(model.names) + list(overlap.names),
            "fixed_Z": {"forward_ms": round(t_f2 * 1e3, 3), "backward_ms": round(t_b2 * 1e3, 3),
                        "bwd_fwd": round(t_b2 / t_f2, 3)}}


def c1_leg(reps: int = 5) -> dict:
    """The plain-DOS graphene configuration at its own size: 64 x 64
    rectangular cells (16,384 atoms, p_z), t = -2.7 eV nearest neighbours,
    onsite 0 stored, window b = 0, a = 8.1 / 0.99, R = 16 Rademacher samples
    (seed 20240917), M = 1024 doubled moments, Jackson DOS on 2001 points;
    FP64 and FP32. Wall time per run against the summed kernel time of the
    same run (the launch gaps of a small operator)."""
    import numpy as np

    from rsatb_amd import _lib
    from rsatb_amd.lattice import graphene_rect_supercell, neighbor_edges
    from rsatb_amd.models import SkPzModel, assemble

    st = graphene_rect_supercell(64, 64)
    H = assemble(st, neighbor_edges(st, 1.6), SkPzModel())
    N, R, M = st.n_atoms, 16, 1024
    x = -1.0 + np.arange(2001) / 1000.0
    out = {"workload": f"graphene 64x64 rectangular cells ({N} atoms, p_z), t=-2.7 eV, window b=0 a=8.1/0.99, "
                       f"R={R} Rademacher (seed 20240917), M={M} doubled moments, Jackson DOS on 2001 points"}
    with _lib.Context(0) as c:
        c.upload_bsr(_lib.SLOT_H, H)
        c.set_operator(_lib.FORM_ORTHOGONAL, 0.0, 8.1 / 0.99)
        c.set_samples(N, R, _lib.SAMPLES_RADEMACHER, 20240917)
        mus = {}
        for name, prec in (("fp64", _lib.PREC_FP64), ("fp32", _lib.PREC_FP32)):
            c.kpm_moments(M, prec)  # warm-up
            t0 = time.perf_counter()
            for _ in range(reps):
                mus[name] = c.kpm_moments(M, prec)
            dt = (time.perf_counter() - t0) / reps
            c.set_profiling(True)
            c.kpm_moments(M, prec)
            ks = {nm: c.kernel_stats(k) for k, nm in _lib.KCLASS_NAMES.items()}
            c.set_profiling(False)
            kern = sum(v["avg_ms"] * v["count"] for v in ks.values()) * 1e-3
            out[name] = {"ms": round(dt * 1e3, 3), "value": float(N) * R * M / dt, "unit": UNIT,
                         "kernel_ms": round(kern * 1e3, 3), "wall_over_kernel": round(dt / kern, 3),
                         "launches": int(sum(v["count"] for v in ks.values()))}
    rho = _lib.kpm_density(mus["fp64"], x, 0.5 / 1000.0)
    out["dos_err_fp32"] = dos_error(mus["fp64"], mus["fp32"])
    out["dos_max"] = float(np.max(2.0 * rho))
    return out


def c2_leg(reps: int) -> dict:
    """Diamond 8 x 8 x 8 conventional cells (53,248 orbitals), sandwich,
    R = 64 random-phase complex vectors, M = 2048: complete runs in FP64 and
    in the FP32 form, with the roofline of the FP64 run's dominant kernel."""
    import numpy as np

    from rsatb_amd import _lib
    from rsatb_amd.driver import RSATB, PreconditionerOptions, SetupOptions
    from rsatb_amd.lattice import diamond_supercell
    from rsatb_amd.models import RandomBlockModel

    st = diamond_supercell(8)
    calc = RSATB(st, RandomBlockModel(), SetupOptions(
        nsamples=64, sample_seed=20240917, compute_bounds=True, sample_kind="phase",
        preconditioner=PreconditionerOptions(operator_form="sandwich")))
    calc._bind_kpoint(0)
    ctx = calc.context
    rows, R, M = st.n_atoms * 13, 64, 2048
    out = {"workload": f"diamond 8x8x8 conventional cells ({st.n_atoms} atoms x 13 orbitals), sandwich Z^H H Z, "
                       f"R={R} random-phase complex vectors, M={M} moments"}
    mus = {}
    for name, prec in (("fp64", _lib.PREC_FP64), ("fp32", _lib.PREC_FP32)):
        ctx.kpm_moments(M, prec)
        t0 = time.perf_counter()
        for _ in range(reps):
            mus[name] = ctx.kpm_moments(M, prec)
        dt = (time.perf_counter() - t0) / reps
        out[name] = {"ms": round(dt * 1e3, 2), "value": rows * R * M / dt}
        if name == "fp64":
            ctx.set_profiling(True)
            ctx.kpm_moments(M, prec)
            stats = {}
            for k, nm in _lib.KCLASS_NAMES.items():
                st_ = ctx.kernel_stats(k)
                st_["name"] = nm
                stats[nm] = st_
            ctx.set_profiling(False)
            out["roofline"] = roofline(stats, "fp64", "diamond8_R64_sandwich")
    out["fp32_speedup"] = round(out["fp32"]["value"] / out["fp64"]["value"], 3)
    out["fp32_max_abs_dmu"] = float(np.max(np.abs(mus["fp32"] - mus["fp64"])))
    out["fp32_dos_err"] = dos_error(mus["fp64"], mus["fp32"])
    ctx.close()
    return out


def sk_silicon_params():
    """(basis, H, S) of a synthetic sp3d5 silicon model, every SK channel
    (s, p, d shells), cutoff 4.1 A (first and second neighbour shells of
    diamond Si), half the channels natural splines and half exp_poly."""
    import numpy as np

    from rsatb_amd.sktb import Radial, SkParameters

    h, s = SkParameters(), SkParameters()
    h.onsite["Si"] = {"s": -4.2, "p": 1.7, "d": 6.5}
    vals = [("ss", "sigma", -1.9), ("sp", "sigma", 2.1), ("pp", "sigma", 3.0), ("pp", "pi", -0.9),
            ("sd", "sigma", -1.6), ("pd", "sigma", -1.3), ("pd", "pi", 1.9), ("dd", "sigma", -1.2),
            ("dd", "pi", 2.2), ("dd", "delta", -1.7)]
    r = np.linspace(2.0, 4.1, 10)
    for i, (sp, bond, v) in enumerate(vals):
        if i % 2:
            h.set_pair_channel("Si", "Si", sp, bond, Radial.exp_poly([v, 0.1 * v, -0.05 * v], 4.1, 0.4, 3.4))
        else:
            h.set_pair_channel("Si", "Si", sp, bond, Radial.spline(r, v * np.exp(-(r - 2.35)), 4.1))
        s.set_pair_channel("Si", "Si", sp, bond, Radial.spline(r, 0.03 * v * np.exp(-(r - 2.35)), 4.1))
    return {"Si": ["s", "p", "d"]}, h, s


def sk_leg(cells: int, reps: int, cpu_edges: int) -> dict:
    """Slater-Koster assembly of H and S on the device (rg_sk_assemble) for
    diamond Si, cells^3 conventional cells, 9 orbitals, 16 neighbours per
    atom; the CPU oracle timed on the first cpu_edges edge blocks."""
    import numpy as np

    from oracle import sk_oracle as so
    from rsatb_amd import _lib
    from rsatb_amd.lattice import diamond_supercell, neighbor_edges
    from rsatb_amd.sktb import SktbModel

    st = diamond_supercell(cells, a=5.43)
    model = SktbModel(*sk_silicon_params())
    e = neighbor_edges(st, model.cutoff)
    pairs = np.stack([e.src, e.dst], axis=1)
    with _lib.Context(0) as c:
        args = (st.atom_type, st.positions, st.cell, pairs, e.shift)
        c.sk_assemble(_lib.SLOT_H, model.arrays("H"), *args)  # warm-up
        best = float("inf")
        for _ in range(reps):
            t0 = time.perf_counter()
            nh = c.sk_assemble(_lib.SLOT_H, model.arrays("H"), *args)
            ns = c.sk_assemble(_lib.SLOT_S, model.arrays("S"), *args)
            best = min(best, time.perf_counter() - t0)
    n_edges = int(e.src.size)
    t0 = time.perf_counter()
    for k in range(cpu_edges):
        so.pair_block(model.table, "H", 0, 0, e.vec[k])
    dt = time.perf_counter() - t0
    gpu_rate = 2 * n_edges / best
    cpu_rate = cpu_edges / dt
    return {"workload": f"diamond Si {cells}^3 conventional cells ({st.n_atoms} atoms x 9 orbitals, s p d), "
                        f"{n_edges} edges (16 neighbours), H and S, Gamma",
            "gpu_ms_H_and_S": round(best * 1e3, 3), "stored_blocks": [nh, ns],
            "value": gpu_rate, "unit": "SK edge blocks/s",
            "cpu_baseline": {"value": cpu_rate, "unit": "SK edge blocks/s", "cores": 1, "kind": "port",
                             "sample": f"first {cpu_edges} edges of H, Python oracle, {dt:.2f} s"},
            "gpu_over_cpu": round(gpu_rate / cpu_rate, 1)}


def kg_leg(cells: int, R: int, M: int, cpu_M: int) -> dict:
    """Kubo-Greenwood spectrum (rg_sk_velocity + rg_kubo_greenwood) on
    graphene, cells x cells rectangular cells, one p_z orbital, nearest-
    neighbour SK hopping; R Rademacher samples, M moments (Jackson). The CPU
    leg runs the oracle (scipy sparse sequences, FP64) with cpu_M moments and
    one sample."""
    import numpy as np

    from oracle import kg_oracle as kg
    from rsatb_amd import _lib
    from rsatb_amd.lattice import graphene_rect_supercell, neighbor_edges
    from rsatb_amd.sktb import Radial, SkParameters, SktbModel

    h = SkParameters().set_onsite("C", "pz", 0.0)
    h.set_pair_channel("C", "C", "pp", "pi", Radial.spline([0.0, 1.6], [-2.7, -2.7], 1.6))
    model = SktbModel({"C": ["pz"]}, h)
    st = graphene_rect_supercell(cells, cells)
    e = neighbor_edges(st, model.cutoff)
    pairs = np.stack([e.src, e.dst], axis=1)
    N = st.n_atoms
    with _lib.Context(0) as c:
        args = (st.atom_type, st.positions, st.cell, pairs, e.shift)
        c.sk_assemble(_lib.SLOT_H, model.arrays("H"), *args)
        c.sk_velocity(model.arrays("H"), *args)
        c.set_operator(_lib.FORM_ORTHOGONAL, 0.0, 8.2)
        c.set_samples(N, R, _lib.SAMPLES_RADEMACHER, 20240917)
        c.kubo_greenwood(16)  # warm-up
        t0 = time.perf_counter()
        sigma, energy = c.kubo_greenwood(M)
        dt = time.perf_counter() - t0
        # per-kernel durations from a second, event-timed pass (timing events
        # idle the GPU a few us at every recorded launch)
        c.set_profiling(True)
        c.kubo_greenwood(M)
        ks = {name: c.kernel_stats(k) for k, name in _lib.KCLASS_NAMES.items()}
        c.set_profiling(False)
        # the same operator for the CPU leg (b = 1: one scalar per block)
        Hb = c.download_bsr(_lib.SLOT_H)
        Bb = [c.download_bsr(sl) for sl in (_lib.SLOT_VX, _lib.SLOT_VY, _lib.SLOT_VZ)]
    gemm = ks["kg_gemm"]
    g_tf = gemm["flops"] / (gemm["avg_ms"] * 1e-3) / 1e12 if gemm["count"] else 0.0
    # CPU: the oracle's sparse sequences and moment matrix for one sample
    import scipy.sparse as sp

    def csr(b, scale=1.0):
        return sp.csr_matrix((scale * b.values.reshape(-1), b.col_ind, b.row_ptr), shape=(N, N))

    Hs = csr(Hb)
    Vs = [csr(b, 1j) for b in Bb]
    r = np.random.default_rng(1).choice([-1.0, 1.0], size=N) / np.sqrt(N)
    t1 = time.perf_counter()
    L, Rq = kg.sequences(Hs, Vs, r, 0.0, 8.2, cpu_M)
    kg.greenwood_from_moments(kg.moment_matrix(L, Rq, cpu_M), cpu_M)
    ct = time.perf_counter() - t1
    units = float(N) * R * M
    return {"workload": f"graphene {cells}x{cells} rectangular cells ({N} atoms, p_z), nearest-neighbour SK, "
                        f"R={R} Rademacher samples, M={M} moments, Greenwood xx/yy/zz, FP64",
            "ms": round(dt * 1e3, 2), "value": units / dt, "unit": "orbital*vector*moment/s",
            "kernels": {k: {"avg_ms": round(v["avg_ms"], 4), "count": v["count"]} for k, v in ks.items()
                        if v["count"]},
            "gemm_TFLOPs": round(g_tf, 2), "gemm_frac_of_fp64_peak": round(g_tf / 78.6, 3),
            "sigma_xx_max": float(np.max(sigma[:, 0])),
            "cpu_baseline": {"value": float(N) * cpu_M / ct, "unit": "orbital*vector*moment/s", "cores": 1,
                             "kind": "port",
                             "sample": f"one sample, {cpu_M} moments, scipy sparse oracle, {ct:.2f} s"}}


def graphene_scale_leg(cells: int, R: int = 16, M: int = 1024) -> dict:
    """The plain-DOS graphene operator (p_z, nearest-neighbour t = -2.7 eV,
    onsite 0 stored, window b = 0, a = 8.1 / 0.99) at cells x cells
    rectangular cells, R Rademacher samples, M doubled moments, FP64 and
    FP32: the b = 1 fused step's achieved HBM bandwidth on a working set that
    does not fit the caches (at 512 x 512 cells: 1,048,576 orbitals, 134 MB
    per FP64 vector)."""
    import numpy as np

    from rsatb_amd import _lib
    from rsatb_amd.lattice import graphene_rect_supercell, neighbor_edges
    from rsatb_amd.sktb import Radial, SkParameters, SktbModel

    h = SkParameters().set_onsite("C", "pz", 0.0)
    h.set_pair_channel("C", "C", "pp", "pi", Radial.spline([0.0, 1.6], [-2.7, -2.7], 1.6))
    model = SktbModel({"C": ["pz"]}, h)
    st = graphene_rect_supercell(cells, cells)
    e = neighbor_edges(st, model.cutoff)
    N = st.n_atoms
    out = {"workload": f"graphene {cells}x{cells} rectangular cells ({N} atoms, p_z), nearest-neighbour t=-2.7 eV, "
                       f"window b=0 a=8.1/0.99, R={R} Rademacher samples, M={M} moments (doubled)"}
    with _lib.Context(0) as c:
        c.sk_assemble(_lib.SLOT_H, model.arrays("H"), st.atom_type, st.positions, st.cell,
                      np.stack([e.src, e.dst], axis=1), e.shift)
        c.set_operator(_lib.FORM_ORTHOGONAL, 0.0, 8.1 / 0.99)
        c.set_samples(N, R, _lib.SAMPLES_RADEMACHER, 20240917)
        mus = {}
        for name, prec in (("fp64", _lib.PREC_FP64), ("fp32", _lib.PREC_FP32)):
            # warm-up at full length: the first run of a length allocates its
            # workspace (the deferred dot partials: 513 MiB here, about 6 ms)
            c.kpm_moments(M, prec, _lib.DOUBLING_ON)
            t0 = time.perf_counter()
            mus[name] = c.kpm_moments(M, prec, _lib.DOUBLING_ON)
            dt = time.perf_counter() - t0
            c.set_profiling(True)  # per-kernel durations: a second, event-timed pass
            c.kpm_moments(M, prec, _lib.DOUBLING_ON)
            st_ = c.kernel_stats(_lib.KCLASS_SPMM_STEP)
            c.set_profiling(False)
            gbs = st_["bytes"] / (st_["avg_ms"] * 1e-3) / 1e9
            out[name] = {"ms": round(dt * 1e3, 2), "value": float(N) * R * M / dt,
                         "unit": "orbital*vector*moment/s", "step_avg_us": round(st_["avg_ms"] * 1e3, 2),
                         "step_algorithmic_MB": round(st_["bytes"] / 1e6, 2), "step_GBs": round(gbs, 1),
                         "step_hbm_frac": round(gbs / 8000.0, 3)}
    out["fp32_speedup"] = round(out["fp64"]["ms"] / out["fp32"]["ms"], 3)
    out["max_abs_dmu_fp32"] = float(np.max(np.abs(mus["fp32"] - mus["fp64"])))
    out["mu0"] = float(mus["fp64"][0].real)
    return out


def tblg_leg(reps: int, R: int, M: int) -> dict:
    """Differentiable Kubo-Greenwood on twisted bilayer graphene: the
    Greenwood spectrum and its derivative with respect to the interlayer
    hopping scale (rg_dc_spectrum_tangent), R Rademacher samples, M moments,
    FP64, window from the Gershgorin bound; the spectrum also in the FP32
    form (FP32 recurrence, FP64 moment GEMMs). sigma(E_F = 0) and its derivative
    are read at the node interpolation of E = 0; the derivative is checked at
    scale against central differences of device spectra (h = 1e-4 and 2e-4,
    Richardson-extrapolated)."""
    import numpy as np

    from rsatb_amd import _lib
    from rsatb_amd.transport import interpolate_spectrum_component

    t_setup = time.perf_counter()
    from rsatb_amd.autograd import tblg_transport_model

    st, e, model = tblg_transport_model(reps)
    N = st.n_atoms
    theta = np.array([1.0, 1.0])
    H = model.hamiltonian.build(theta)
    rows = np.repeat(np.arange(N), np.diff(H.row_ptr))
    half = float(np.max(np.bincount(rows, weights=np.abs(H.values.reshape(-1)), minlength=N))) + 0.05
    volume = float(abs(np.linalg.det(st.cell)))
    pref = np.pi * 2.0 * 4.0 * N / volume * 2.4341348e6  # S/m with both spins, axis "abc"
    e1 = np.array([0.0, 1.0])
    setup_s = time.perf_counter() - t_setup

    def bind(c, th):
        c.upload_bsr(_lib.SLOT_H, model.hamiltonian.build(th))
        for a in range(3):
            c.upload_bsr(_lib.SLOT_VX + a, model.velocity[a].build(th))

    with _lib.Context(0) as c:
        c.set_samples(N, R, _lib.SAMPLES_RADEMACHER, 20240917)
        c.set_operator(_lib.FORM_ORTHOGONAL, 0.0, half)
        bind(c, theta)
        c.upload_bsr(_lib.SLOT_DH, model.hamiltonian.build(e1))
        for a in range(3):
            c.upload_bsr(_lib.SLOT_DVX + a, model.velocity[a].build(e1))
        c.dc_spectrum_tangent(16)  # warm-up
        t0 = time.perf_counter()
        sigma0, energy = c.dc_spectrum(M, 0, _lib.KERNEL_JACKSON, pref)
        dt0 = time.perf_counter() - t0
        c.dc_spectrum(16, 0, _lib.KERNEL_JACKSON, pref, precision=_lib.PREC_FP32)  # warm-up
        t0 = time.perf_counter()
        sigma32, _ = c.dc_spectrum(M, 0, _lib.KERNEL_JACKSON, pref, precision=_lib.PREC_FP32)
        dt32 = time.perf_counter() - t0
        t0 = time.perf_counter()
        sigma, dsigma, energy = c.dc_spectrum_tangent(M, 0, _lib.KERNEL_JACKSON, pref)
        dt = time.perf_counter() - t0
        c.set_profiling(True)  # per-kernel durations: a second, event-timed pass
        c.dc_spectrum_tangent(M, 0, _lib.KERNEL_JACKSON, pref)
        ks = {name: c.kernel_stats(k) for k, name in _lib.KCLASS_NAMES.items()}
        c.set_profiling(False)
        def central(h):
            bind(c, theta + h * e1)
            sp, _ = c.dc_spectrum(M, 0, _lib.KERNEL_JACKSON, pref)
            bind(c, theta - h * e1)
            sm, _ = c.dc_spectrum(M, 0, _lib.KERNEL_JACKSON, pref)
            return (sp - sm) / (2 * h)

        # Richardson: the O(h^2) terms of two central differences cancel
        fd = (4.0 * central(1e-4) - central(2e-4)) / 3.0
    s_f = interpolate_spectrum_component(energy, sigma, 0.0)
    ds_f = interpolate_spectrum_component(energy, dsigma, 0.0)
    fd_f = interpolate_spectrum_component(energy, fd, 0.0)
    diag = [0, 3]
    return {"workload": f"twisted bilayer graphene 1.05 deg ((31,32) cell, 11908 atoms) tiled {reps}x{reps} "
                        f"({N} atoms, p_z, {e.src.size} bonds within 3.8 A), orthogonal, R={R} Rademacher samples, "
                        f"M={M} moments, Greenwood, d/d(interlayer hopping scale), FP64",
            "ms_sigma": round(dt0 * 1e3, 1), "ms_sigma_and_tangent": round(dt * 1e3, 1),
            "ms_sigma_fp32": round(dt32 * 1e3, 1), "fp32_speedup": round(dt0 / dt32, 3),
            "fp32_max_diff_rel": float(np.max(np.abs(sigma32 - sigma0)) / np.max(np.abs(sigma0))),
            "fp32_EF0_rel_diff": float(np.max(np.abs(interpolate_spectrum_component(energy, sigma32, 0.0)[[0, 3]]
                                                     - interpolate_spectrum_component(energy, sigma0, 0.0)[[0, 3]]))
                                       / np.max(np.abs(interpolate_spectrum_component(energy, sigma0, 0.0)[[0, 3]]))),
            "tangent_over_sigma": round(dt / dt0, 2), "host_setup_s": round(setup_s, 1),
            "value": float(N) * R * M / dt, "unit": "orbital*vector*moment/s (spectrum + tangent)",
            "kernels": {k: {"avg_ms": round(v["avg_ms"], 4), "count": v["count"]} for k, v in ks.items()
                        if v["count"]},
            "sigma_EF0_S_per_m": [float(s_f[i]) for i in diag],
            "dsigma_dinter_EF0_S_per_m": [float(ds_f[i]) for i in diag],
            "fd_rel_diff_EF0": float(np.max(np.abs(ds_f[diag] - fd_f[diag])) / np.max(np.abs(fd_f[diag]))),
            "fd_rel_diff_spectrum": float(np.max(np.abs(dsigma - fd)) / np.max(np.abs(fd))),
            "sigma_matches_plain_call": bool(np.max(np.abs(sigma - sigma0)) <= 1e-10 * np.max(np.abs(sigma0)))}


def per_moment_error(mu_ref, mu, blocks: int = 16) -> dict:
    """|mu_n - mu_ref_n| over all moments: its maximum, the moment where it
    occurs, and the maximum within each of `blocks` equal ranges of n."""
    import numpy as np

    d = np.abs(mu - mu_ref)
    edges = np.linspace(0, d.size, blocks + 1).astype(int)
    return {"max": float(np.max(d)), "argmax": int(np.argmax(d)),
            "block_max": [float(np.max(d[edges[k]:edges[k + 1]])) for k in range(blocks)],
            "block_edges": [int(e) for e in edges]}


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

    legs = [("c2", not args.no_c2, lambda: c2_leg(3)),
            ("c1", not args.no_c1, lambda: c1_leg()),
            ("adjoint", not args.no_adjoint, lambda: adjoint_leg(args.adj_cells, max(1, min(args.steps, 5)))),
            ("sk_assembly", not args.no_sk, lambda: sk_leg(args.sk_cells, 3, args.sk_cpu_edges)),
            ("kubo_greenwood", not args.no_kg,
             lambda: kg_leg(args.kg_cells, args.kg_samples, args.kg_moments, args.kg_cpu_moments)),
            ("graphene_scale", not args.no_graphene_scale, lambda: graphene_scale_leg(args.graphene_cells)),
            ("tblg_transport", not args.no_tblg, lambda: tblg_leg(args.tblg_reps, args.tblg_samples, args.tblg_moments))]
    if args.leg:
        fn = dict((k, f) for k, _, f in legs)[args.leg]
        print(json.dumps({"leg": args.leg, args.leg: fn()}), flush=True)
        return 0

    t_setup = time.perf_counter()
    st = diamond_supercell(args.cells)
    calc = RSATB(st, RandomBlockModel(), SetupOptions(
        nsamples=args.nsamples, sample_seed=20240917, compute_bounds=True, sample_kind="phase",
        preconditioner=PreconditionerOptions(operator_form="sandwich")))
    calc._bind_kpoint(0)
    ctx = calc.context
    setup_s = time.perf_counter() - t_setup
    rows = st.n_atoms * 13
    R, Ms = args.nsamples, args.step_moments
    log(f"setup {setup_s:.1f}s: {rows} orbitals, bounds {calc.energy_bounds()}, "
        f"patches {calc._kstates[0].get('isqrt')}")
    prec64 = _lib.PREC_FP32 if args.fp32_only else _lib.PREC_FP64
    if args.fp32_only:
        args.no_fp32 = True

    for _ in range(args.warmup):
        ctx.kpm_moments(Ms, prec64)
    log(f"warm-up done ({args.warmup} steps)")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        mu_step = ctx.kpm_moments(Ms, prec64)  # synchronous on return
    dt = time.perf_counter() - t0
    log(f"timed FP64 steps: {dt / args.steps * 1e3:.1f} ms per step")
    # The same K steps again with a HIP event pair around every launch on the
    # library's stream: the per-kernel durations behind the roofline. This
    # pass is timed separately, because the events idle the GPU a few us at
    # every recorded launch.
    ctx.set_profiling(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.kpm_moments(Ms, prec64)
    dt_prof = time.perf_counter() - t0
    stats = {}
    for k, name in _lib.KCLASS_NAMES.items():
        s_ = ctx.kernel_stats(k)
        s_["name"] = name
        stats[name] = s_
    ctx.set_profiling(False)
    ms = dt / args.steps * 1e3
    value = rows * R * Ms / (dt / args.steps)
    tag = f"diamond{args.cells}_R{R}_sandwich"

    line = {
        "metric": METRIC, "value": value, "unit": UNIT, "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "ms_per_step_event_timed": dt_prof / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32" if args.fp32_only else "f64", "data": "synthetic",
        "config": {"workload": f"diamond {args.cells}x{args.cells}x{args.cells} conventional cells "
                               f"({st.n_atoms} atoms x 13 orbitals = {rows} orbitals), sandwich Z^H H Z with "
                               f"graph-local Z = S^-1/2, R={R} random-phase complex vectors, M={args.moments} "
                               f"moments; one step = the first {Ms} moments ({Ms // 2} doubled sandwich applies) "
                               f"at full N and R",
                   "n_orbitals": rows, "R": R, "M": args.moments, "step_moments": Ms, "form": "sandwich",
                   "precision": "fp32" if args.fp32_only else "fp64", "bounds": list(calc.energy_bounds()),
                   "setup_s": round(setup_s, 1)},
        "roofline": roofline(stats, "fp32" if args.fp32_only else "fp64", tag),
        "kernels": {k: {"avg_ms": round(v["avg_ms"], 5), "count": v["count"]} for k, v in stats.items()},
        "mu0": float(mu_step[0].real),
    }
    if not args.no_fp32:
        ctx.kpm_moments(Ms, _lib.PREC_FP32)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ctx.kpm_moments(Ms, _lib.PREC_FP32)
        dt32 = time.perf_counter() - t0
        v32 = rows * R * Ms / (dt32 / args.steps)
        line["fp32"] = {"value": v32, "ms_per_step": dt32 / args.steps * 1e3, "speedup_vs_fp64": v32 / value,
                        "dtype": "f32 MFMA products and recurrence, f64 dots"}
        log(f"timed FP32 steps: {dt32 / args.steps * 1e3:.1f} ms per step")
    mu64 = None
    if not args.no_full:
        M = args.moments
        t0 = time.perf_counter()
        mu64 = ctx.kpm_moments(M, prec64)
        t64 = time.perf_counter() - t0
        full = {"fp64_s": round(t64, 3), "fp64_value": rows * R * M / t64}
        log(f"complete FP64 run (M={M}): {t64:.2f} s")
        if not args.no_fp32:
            t0 = time.perf_counter()
            mu32 = ctx.kpm_moments(M, _lib.PREC_FP32)
            t32 = time.perf_counter() - t0
            log(f"complete FP32 run (M={M}): {t32:.2f} s")
            full.update({"fp32_s": round(t32, 3), "fp32_value": rows * R * M / t32,
                         "fp32_speedup": round(t64 / t32, 3), "dos_err": dos_error(mu64, mu32),
                         "dos_err_definition": "max_E |rho32 - rho64| / max_E rho64, Jackson, 2001 points",
                         "per_moment_abs_err": per_moment_error(mu64, mu32)})
        line["full_run"] = full
        line["mu0_full"] = float(mu64[0].real)
        line["max_abs_mu"] = float(np.max(np.abs(mu64)))
    if not args.no_cpu_baseline:
        ref = mu64 if mu64 is not None else mu_step
        line["cpu_baseline"] = cpu_baseline(calc, min(args.cpu_moments, len(ref)), ref)
    ctx.close()
    del calc
    if not args.no_legs:
        for key, on, fn in legs:
            if on:
                t0 = time.perf_counter()
                line[key] = fn()
                log(f"leg {key}: {time.perf_counter() - t0:.1f} s")
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
