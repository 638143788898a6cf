"""Benchmark of the OpenCAL hot path on one MI355X.

Headline workload: SciddicaT debris flow, FP64, 4096 x 4096 toroidal grid,
full grid (CAL_NO_OPT), driven through OpenCAL's C API (calRunCAStep2D per
step: global transition on the GPU, host steering callback, implicit update).
Synthetic input: 500 m slope + 20 Gaussian bumps, 40 x 40 source patch.

value = cell updates per second (GCUPS) over the timed steps with the model
already resident in HBM.  The same JSON line carries the roofline of the
fused sweep kernel (HIP events on the model's stream), the CPU baseline (the
reference OpenCAL-OMP build when present, else the oracle restatement) timed
on this host, and the secondary workloads (active cells, Life, batched
ensemble, adjoint) when --all is given.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--all]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL_SCIDDICAT = 24.0   # z read, h read, h written (FP64)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def bench_sciddicat_plain(n, steps, warmup, seed=2024):
    from opencal_amd import synth
    from opencal_amd.models import SciddicaT
    z, h, _ = synth.sciddicat_inputs(n, n, seed)
    m = SciddicaT(z, h, 0.5, 0.001, steps + warmup, active=False)
    m.init()
    m.steps(warmup)          # first step uploads the model
    m.sync()
    m.timing(True)
    t0 = time.perf_counter()
    m.steps(steps)
    m.sync()
    wall = time.perf_counter() - t0
    ktime_ms, launches = m.kernel_time()
    m.timing(False)
    st = m.stats()
    _, hh, _ = m.read()
    mass_err = abs(hh.sum() - h.sum()) / h.sum()
    m.close()
    cells = n * n
    return {
        "wall_s": wall, "gcups": cells * steps / wall / 1e9,
        "kernel_ms_avg": ktime_ms / max(launches, 1), "kernel_launches": launches,
        "achieved_gbs": BYTES_PER_CELL_SCIDDICAT * cells / (ktime_ms / max(launches, 1) * 1e-3) / 1e9,
        "launches_per_step": st["launches"] / max(st["steps"], 1), "mass_rel_err": mass_err,
    }


def cpu_baseline_sciddicat(n, steps, seed=2024):
    """Reference OpenCAL-OMP (oracle/_ref build) or the oracle restatement,
    full grid, timed on this host's cores."""
    import native
    from opencal_amd import synth
    z, h, _ = synth.sciddicat_inputs(n, n, seed)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    kind = "reference" if native.ref("sciddicat") is not None else "port"
    fn = native.ref_sciddicat if kind == "reference" else native.oracle_sciddicat
    fn(z[:64, :64].copy(), h[:64, :64].copy(), 0.5, 0.001, 2)   # warm-up
    t0 = time.perf_counter()
    fn(z, h, 0.5, 0.001, steps)
    wall = time.perf_counter() - t0
    return {"value": n * n * steps / wall / 1e9, "unit": "GCUPS", "cores": threads, "kind": kind,
            "sample": f"SciddicaT {n}x{n} full grid, {steps} steps (incl. init), OMP_NUM_THREADS={threads}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--size", type=int, default=4096)
    ap.add_argument("--cpu-steps", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--all", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the engine runs on one MI355X, "
              "multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)

    import torch
    assert torch.cuda.is_available(), "needs a GPU"
    log(f"sciddicaT plain {args.size}^2: warmup {args.warmup}, timed {args.steps}")
    r = bench_sciddicat_plain(args.size, args.steps, args.warmup)
    log(json.dumps(r))
    out = {
        "metric": "SciddicaT FP64 cell updates per second, full grid (GCUPS)",
        "value": round(r["gcups"], 3),
        "unit": "GCUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["wall_s"] / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded DEM: 500 m slope + 20 Gaussian bumps, 40x40 source patch U(1,10) m)",
        "config": {"workload": f"sciddicaT_plain_{args.size}x{args.size}_fp64", "rows": args.size,
                   "cols": args.size, "p_r": 0.5, "p_epsilon": 0.001, "optimization": "CAL_NO_OPT",
                   "driver": "calRunCAStep2D per step (C API)"},
        "roofline": {"bound": "hbm", "achieved": round(r["achieved_gbs"], 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(r["achieved_gbs"] / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "step_fused", "kernel_us": round(r["kernel_ms_avg"] * 1e3, 2),
                     "bytes_per_cell": BYTES_PER_CELL_SCIDDICAT},
    }
    if not args.no_cpu:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline_sciddicat(args.size, args.cpu_steps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
