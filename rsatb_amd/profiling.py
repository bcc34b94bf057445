"""Summaries of rocprofv3 output for this library's kernels.

Turns three rocprofv3 runs of the same command into a short text summary and
a JSON record of per-launch HBM traffic:

  1. --kernel-trace --stats            per-kernel calls and average duration
  2. --pmc FETCH_SIZE                  KiB read from HBM per dispatch
  3. --pmc WRITE_SIZE                  KiB written to HBM per dispatch

The counters are scaled by what they read for a plain copy kernel of known
size in the same run (k_stage, which copies the sample block: `copy_bytes`
read, and `copy_write_bytes` written, the same unless it narrows to FP32), so
traffic is in bytes:
traffic = FETCH_SIZE * (copy_bytes / FETCH_SIZE(copy)) + WRITE_SIZE *
(copy_write_bytes / WRITE_SIZE(copy)).

    python -m rsatb_amd.profiling --trace D1 --fetch D2 --write D3 \\
        --copy-bytes N --command "..." --title "..." --tag T --out profiles/x.md --json profiles/x.json
"""
from __future__ import annotations

import argparse
import csv
import json
import re
from collections import defaultdict
from pathlib import Path

# kernel-class labels of rg_kernel_stats, from the MFMA SpMM template's mode
_MODE_CLASS = {"0": "spmm_plain", "1": "spmm_step", "2": "spmm_step", "3": "spmm_step", "4": "spmm_adjoint"}


def _csv(path: Path):
    with open(path, newline="") as f:
        return list(csv.DictReader(f))


def _find(d: Path, suffix: str) -> Path:
    hits = sorted(Path(d).rglob(f"*{suffix}"))
    if not hits:
        raise FileNotFoundError(f"no *{suffix} under {d}")
    return hits[0]


def short_name(name: str) -> str:
    """'void rg::(anonymous namespace)::k_spmm_dma<double, 8, 2, true, 2>(...)' -> 'k_spmm_dma<double, 8, 2, true, 2>'."""
    m = re.search(r"(k_\w+(?:<[^()]*>)?)\s*\(", name)
    return m.group(1) if m else name.split("(")[0]


def kernel_class(short: str) -> tuple[str, str] | None:
    """(class, dtype) of an SpMM / reduction kernel, as rg_kernel_stats labels them."""
    m = re.match(r"k_spmm_(?:mfma|dma|grp)<(\w+), \d+, (\d)", short)
    if m:
        return _MODE_CLASS.get(m.group(2), "spmm"), "f64" if m.group(1) == "double" else "f32"
    if short.startswith("k_reduce"):
        return "reduce", "f64"
    if short.startswith("k_group"):
        return "group_dots", "f64"
    return None


def read_stats(d: Path):
    rows = _csv(_find(d, "kernel_stats.csv"))
    out = []
    for r in rows:
        out.append({"kernel": short_name(r["Name"]), "calls": int(r["Calls"]),
                    "avg_us": float(r["AverageNs"]) / 1e3, "pct": float(r["Percentage"])})
    return out


def read_pmc(d: Path, counter: str):
    """Mean counter value per dispatch, per kernel."""
    acc = defaultdict(list)
    for r in _csv(_find(d, "counter_collection.csv")):
        if r["Counter_Name"] == counter:
            acc[short_name(r["Kernel_Name"])].append(float(r["Counter_Value"]))
    return {k: sum(v) / len(v) for k, v in acc.items()}


def summarize(trace: Path, fetch: Path, write: Path, copy_bytes: float, copy_write_bytes: float | None = None,
              copy_kernel: str = "k_stage"):
    if copy_write_bytes is None:
        copy_write_bytes = copy_bytes
    stats = read_stats(trace)
    f = read_pmc(fetch, "FETCH_SIZE")
    w = read_pmc(write, "WRITE_SIZE")
    if copy_kernel not in f or copy_kernel not in w:
        raise RuntimeError(f"copy kernel {copy_kernel} missing from the counter runs")
    f_scale = copy_bytes / (f[copy_kernel] * 1024.0)
    w_scale = copy_write_bytes / (w[copy_kernel] * 1024.0)
    kernels = {}
    for s in stats:
        k = s["kernel"]
        rec = dict(s)
        if k in f and k in w:
            rec["fetch_bytes"] = f[k] * 1024.0 * f_scale
            rec["write_bytes"] = w[k] * 1024.0 * w_scale
            rec["traffic_bytes"] = rec["fetch_bytes"] + rec["write_bytes"]
        cls = kernel_class(k)
        if cls:
            rec["class"], rec["dtype"] = cls
        kernels[k] = rec
    calib = {"copy_kernel": copy_kernel, "copy_bytes": copy_bytes, "copy_write_bytes": copy_write_bytes,
             "fetch_kib_read": f[copy_kernel], "write_kib_read": w[copy_kernel],
             "fetch_scale": f_scale, "write_scale": w_scale}
    return {"calibration": calib, "kernels": kernels}


def render(summary, title: str, commands: list[str]) -> str:
    c = summary["calibration"]
    lines = [f"# {title}", "", "Commands (from the repository root on the GPU box; outputs under gpurun_out/):", ""]
    lines += [f"    {cmd}" for cmd in commands]
    lines += ["", f"Calibration: {c['copy_kernel']} reads {c['copy_bytes'] / 1e6:.1f} MB and writes "
              f"{c['copy_write_bytes'] / 1e6:.1f} MB; FETCH_SIZE reads "
              f"{c['fetch_kib_read'] * 1024 / 1e6:.1f} MB (scale {c['fetch_scale']:.3f}), WRITE_SIZE reads "
              f"{c['write_kib_read'] * 1024 / 1e6:.1f} MB (scale {c['write_scale']:.3f}).", "",
              "| kernel | calls | avg us | % time | HBM read MB | HBM write MB | traffic MB |",
              "|---|---|---|---|---|---|---|"]
    for k, r in sorted(summary["kernels"].items(), key=lambda kv: -kv[1]["pct"]):
        if r["pct"] < 0.05:
            continue
        rd = f"{r['fetch_bytes'] / 1e6:.1f}" if "fetch_bytes" in r else "-"
        wr = f"{r['write_bytes'] / 1e6:.1f}" if "write_bytes" in r else "-"
        tr = f"{r['traffic_bytes'] / 1e6:.1f}" if "traffic_bytes" in r else "-"
        lines.append(f"| `{k}` | {r['calls']} | {r['avg_us']:.1f} | {r['pct']:.1f} | {rd} | {wr} | {tr} |")
    return "\n".join(lines) + "\n"


def read_pmc_all(d: Path):
    """{kernel: {counter: mean value per dispatch}} of one --pmc run."""
    acc = defaultdict(lambda: defaultdict(list))
    for r in _csv(_find(d, "counter_collection.csv")):
        acc[short_name(r["Kernel_Name"])][r["Counter_Name"]].append(float(r["Counter_Value"]))
    return {k: {c: sum(v) / len(v) for c, v in cs.items()} for k, cs in acc.items()}


def render_counters(d: Path, title: str, commands: list[str], simds: int, kernels: list[str],
                    grbm_instances: int = 8) -> str:
    """Per-kernel L2 hit rate (TCC_HIT_sum / (TCC_HIT_sum + TCC_MISS_sum)) and
    MFMA utilisation (SQ_VALU_MFMA_BUSY_CYCLES / (active cycles x SIMDs)); the
    CSV's GRBM_GUI_ACTIVE is summed over the XCDs' instances, so the active
    cycles are GRBM_GUI_ACTIVE / grbm_instances."""
    pm = read_pmc_all(d)
    lines = [f"# {title}", "", "Command (from the repository root on the GPU box; output under gpurun_out/):", ""]
    lines += [f"    {cmd}" for cmd in commands]
    lines += ["", f"MFMA utilisation = SQ_VALU_MFMA_BUSY_CYCLES / (GRBM_GUI_ACTIVE / {grbm_instances} XCDs x "
              f"{simds} SIMDs); "
              "L2 hit rate = TCC_HIT_sum / (TCC_HIT_sum + TCC_MISS_sum); means per dispatch.", "",
              "| kernel | L2 hit rate | L2 misses per dispatch | MFMA utilisation |",
              "|---|---|---|---|"]
    for k in kernels:
        c = pm.get(k)
        if not c:
            continue
        hit, miss = c.get("TCC_HIT_sum", 0.0), c.get("TCC_MISS_sum", 0.0)
        busy, active = c.get("SQ_VALU_MFMA_BUSY_CYCLES"), c.get("GRBM_GUI_ACTIVE")
        hr = f"{hit / (hit + miss):.3f}" if hit + miss > 0 else "-"
        mu = f"{busy / (active / grbm_instances * simds):.3f}" if busy is not None and active else "-"
        lines.append(f"| `{k}` | {hr} | {miss:.3e} | {mu} |")
    return "\n".join(lines) + "\n"


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--counters", type=Path, default=None,
                    help="summarize one --pmc run of TCC_HIT_sum TCC_MISS_sum SQ_VALU_MFMA_BUSY_CYCLES "
                         "GRBM_GUI_ACTIVE instead (with --simds, --kernel)")
    ap.add_argument("--simds", type=int, default=1024)
    ap.add_argument("--grbm-instances", type=int, default=8, help="XCDs the GRBM_GUI_ACTIVE value is summed over")
    ap.add_argument("--kernel", action="append", default=[])
    ap.add_argument("--trace", type=Path)
    ap.add_argument("--fetch", type=Path)
    ap.add_argument("--write", type=Path)
    ap.add_argument("--copy-bytes", type=float)
    ap.add_argument("--copy-write-bytes", type=float, default=None)
    ap.add_argument("--title", required=True)
    ap.add_argument("--tag", default="", help="workload tag the benchmark matches, e.g. diamond8_R64_sandwich")
    ap.add_argument("--command", action="append", default=[])
    ap.add_argument("--out", required=True, type=Path)
    ap.add_argument("--json", type=Path)
    a = ap.parse_args(argv)
    if a.counters is not None:
        a.out.write_text(render_counters(a.counters, a.title, a.command, a.simds, a.kernel, a.grbm_instances))
        return
    s = summarize(a.trace, a.fetch, a.write, a.copy_bytes, a.copy_write_bytes)
    a.out.write_text(render(s, a.title, a.command))
    s["title"] = a.title
    s["workload_tag"] = a.tag
    s["commands"] = a.command
    a.json.write_text(json.dumps(s, indent=1, sort_keys=True) + "\n")


if __name__ == "__main__":
    main()
