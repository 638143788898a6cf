"""Benchmark of the OpenCAL hot path on one MI355X.

Headline (`value`): SciddicaT debris flow, FP64, 4096 x 4096 toroidal grid,
full grid (CAL_NO_OPT), driven through OpenCAL's C API as the reference
example drives it: the steps of calRun2D's loop (calhipRunSteps2D: global
transition, the model's steering — the reference flow reset, declared with
calhipBindSciddicaTSteering2D so that its device twin runs — and the implicit
update), with the model already resident in HBM.  Between two steps the
flows are reset, so a step is a function of z and h alone and the engine
advances the grid up to four steps per launch: the tiles within reach of a
wet cell in LDS; every other tile stays as it is (h + 0.0), which needs no
work once both layers hold it (two consecutive copies).  The same workload
with one calRunCAStep2D call per step (which the engine groups four to a
launch) is reported beside it (`per_step_driver`).  Unit: giga cell
updates per second.

The same JSON line carries
  roofline      the full-grid kernel's measured HBM bytes per launch (the
                committed rocprofv3 counter summary) over its average launch
                time (HIP events on the model's stream), against the 8 TB/s
                HBM peak, beside the launch's minimal bytes (the listed
                tiles' z and h read, h written);
  cpu_baseline  the reference OpenCAL-OMP build (oracle/_ref, when present;
                else the oracle restatement) timed on this host's cores on a
                bounded sample of the same workload;
  workloads     the other measured configurations: SciddicaT active cells
                1024^2 x 4000, the same 4096^2 full grid with debris on
                every cell (the minimisation everywhere), Life 512^2 x 1000,
                the batched ensemble of 256 uneven instances x 2000 steps
                against the same instances run
                one after another, and the adjoint gradient / forward time
                ratio at 1024^2 x 2000 with checkpoints every 45 steps, and
                SCIARA-fv2 2048^2 with active cells for 10000 steps; each
                runs in a child process (`bench.py --only NAME`) that this
                run waits for, so that a failure there is reported in its
                record and leaves the headline standing.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--quick]
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

HBM_PEAK_GBS = 8000.0
# The bytes a full-grid SciddicaT launch must move per cell, whether it
# advances one step or two: h read once and written once (FP64).  z (8 B
# more) is needed only where a wet cell is within reach of the launch's
# steps; the fused sweeps read it only there.
MIN_BYTES_PER_CELL_SCIDDICAT = 16.0
TILES_PROFILE = os.path.join(ROOT, "profiles", "sciddicat_tiles_4096x4096.md")


def profiled_traffic(path=TILES_PROFILE, key="traffic_bytes_per_launch"):
    """Bytes per launch of the full-grid tile kernel from the committed
    rocprofv3 counter summary: key traffic_bytes_per_launch (FETCH_SIZE and
    WRITE_SIZE, calibrated against a device copy) or
    algorithmic_bytes_per_launch (the listed tiles' z and h read, h written);
    None when the summary is absent."""
    try:
        for line in open(path):
            if line.startswith(key + ":"):
                return int(line.split(":")[1])
    except OSError:
        pass
    return None


def tiles_roofline(r, size):
    """The roofline record of the full-grid tile kernel: the measured HBM
    bytes of a launch (committed counter summary, 4096^2 only) over the
    launch's average duration in this run, against the HBM peak; the
    launch's minimal bytes over the same time beside it."""
    kus = r["kernel_ms_avg"] * 1e3
    traffic = profiled_traffic() if size == 4096 else None
    alg = profiled_traffic(key="algorithmic_bytes_per_launch") if size == 4096 else None
    rl = {"bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": None,
          "traffic": traffic,
          "traffic_source": "profiles/sciddicat_tiles_4096x4096.md (rocprofv3 FETCH_SIZE x 2.0 + WRITE_SIZE, "
                            "per 4-step launch, the launches after each model's first two)",
          "kernel": "plain_tiles<4>", "kernel_us": round(kus, 2), "steps_per_launch": round(r["steps_per_launch"], 4),
          "algorithmic_bytes_per_launch": alg,
          "note": "achieved = measured HBM bytes per launch / the launch's average duration (HIP events on the "
                  "model's stream around the timed steps, over their launches); algorithmic_gbs = the launch's "
                  "minimal bytes (the listed tiles' z and h read once, h written once, 24 B per cell; the other "
                  "tiles need nothing once both layers hold them) over the same time.  The launch is bound by "
                  "the listed tiles' dependent steps on one CU each, not by HBM."}
    if traffic:
        rl["achieved"] = round(traffic / (kus * 1e-6) / 1e9, 1)
        rl["frac"] = round(rl["achieved"] / HBM_PEAK_GBS, 4)
        rl["algorithmic_gbs"] = round(alg / (kus * 1e-6) / 1e9, 1)
    return rl


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# --------------------------------------------------------------- headline

def bench_sciddicat_plain(n, steps, warmup, seed=2024, batched=True):
    """batched: the steps of calRun2D's loop in one call (calhipRunSteps2D),
    as the reference example's calRun2D runs them: the full grid advances two
    steps per sweep.  Otherwise one calRunCAStep2D call (one sweep) per step."""
    from opencal_amd import synth
    from opencal_amd.models import SciddicaT
    z, h, _ = synth.sciddicat_inputs(n, n, seed)
    m = SciddicaT(z, h, 0.5, 0.001, steps + warmup + 4, active=False)
    m.init()
    m.steps(warmup)          # the first step uploads the model
    if batched:
        m.run_steps(4)       # the sweep's shape and the tile list are set up on first use
    m.sync()
    m.timing(2)              # one HIP event pair on the model's stream around the timed steps
    t0 = time.perf_counter()
    if batched:
        m.run_steps(steps)
    else:
        m.steps(steps)
    m.sync()
    wall = time.perf_counter() - t0
    ktime_ms, launches = m.kernel_time()
    m.timing(0)
    st = m.stats()
    _, hh, _ = m.read()
    mass_err = abs(hh.sum() - h.sum()) / h.sum()
    m.close()
    cells = n * n
    kavg = ktime_ms / max(launches, 1)
    return {
        "wall_s": wall, "gcups": cells * steps / wall / 1e9, "kernel_ms_avg": kavg, "kernel_launches": launches,
        "steps_per_launch": steps / max(launches, 1),
        "launches_per_step": st["launches"] / max(st["steps"], 1), "mass_rel_err": mass_err,
        "uploads": st["uploads"],
    }


def pcie_inclusive_sciddicat(n, steps, seed=2024):
    """The same workload as one calRun2D call on host arrays: the model's init
    on the host, upload of every substate, the steps, download at the end."""
    from opencal_amd import synth
    from opencal_amd.models import SciddicaT
    z, h, _ = synth.sciddicat_inputs(n, n, seed)
    m = SciddicaT(z, h, 0.5, 0.001, steps, active=False)
    t0 = time.perf_counter()
    m.run()
    m.read()
    wall = time.perf_counter() - t0
    m.close()
    return {"gcups": round(n * n * steps / wall / 1e9, 3), "wall_s": round(wall, 4),
            "note": "calRun2D from host arrays incl. host init, upload of 6 substates x 2 layers and read-back"}


def cpu_record(value, unit, kind, sample):
    """A cpu_baseline object: value in unit, the host threads used (the
    OpenMP setting the reference build runs with), what was timed."""
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": round(value, 6), "unit": unit, "cores": threads, "kind": kind, "sample": sample}


def cpu_baseline_sciddicat(n, steps, seed=2024):
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
    return {"value": round(n * n * steps / wall / 1e9, 5), "unit": "GCUPS", "cores": threads, "kind": kind,
            "sample": f"SciddicaT {n}x{n} full grid, {steps} steps incl. init, OMP_NUM_THREADS={threads}, "
                      f"{wall:.1f} s"}


# ----------------------------------------------------------- other configs

def bench_active(n=1024, steps=4000, seed=7, cpu_steps=400):
    """SciddicaT with CAL_OPT_ACTIVE_CELLS (config 2), stepped through the C
    API two ways: one calRunCAStep2D per step (the reference example's loop,
    CAL_RUN_LOOP with its stop condition; the engine groups four steps to a
    launch) and calhipRunSteps2D with the stop condition removed (the steps
    go to the device program together: 4 steps per launch)."""
    import native
    from opencal_amd import synth
    from opencal_amd.models import SciddicaT
    z, h, _ = synth.sciddicat_inputs(n, n, seed)
    _, oh, counts = native.oracle_sciddicat(z, h, 0.5, 0.001, steps, True, counts=True)
    active_updates = float(np.sum(counts[1:steps].astype(np.float64)))
    res = {}
    for mode in ("steps", "run_steps"):
        # the timed run, then the same run with per-launch HIP events (their
        # records cost host time, so they stay out of the timed one)
        for timed_kernels in (False, True):
            m = SciddicaT(z, h, 0.5, 0.001, steps, active=True)
            m.init()
            if mode == "run_steps":
                m.clear_stop()
            m.steps(1)
            m.sync()
            if timed_kernels:
                m.timing(True)
            t0 = time.perf_counter()
            if mode == "steps":
                m.steps(steps - 1)
            else:
                m.run_steps(steps - 1)
            m.sync()
            wall = time.perf_counter() - t0
            if timed_kernels:
                ktime_ms, launches = m.kernel_time()
                m.timing(False)
            _, gh, count = m.read()
            m.close()
            if not timed_kernels:
                res[mode] = {"ms_per_step": wall / (steps - 1) * 1e3, "active_gcups": active_updates / wall / 1e9,
                             "bit_exact_vs_oracle": bool(np.array_equal(gh, oh) and count == counts[-1])}
            else:
                res[mode]["launches_per_step"] = launches / (steps - 1)
                res[mode]["kernel_us_per_launch"] = ktime_ms / max(launches, 1) * 1e3
                res[mode]["bit_exact_vs_oracle"] &= bool(np.array_equal(gh, oh) and count == counts[-1])
    r1, rk = res["steps"], res["run_steps"]
    out = {"workload": f"sciddicaT_active_{n}x{n}_x{steps}", "ms_per_step": r1["ms_per_step"],
           "active_gcups": r1["active_gcups"], "full_grid_equiv_gcups": n * n * (steps - 1) / (r1["ms_per_step"] * 1e-3 * (steps - 1)) / 1e9,
           "driver": "calRunCAStep2D per step (CAL_RUN_LOOP with the stop condition, as the reference example)",
           "per_step": r1, "run_steps": rk,
           "final_active_cells": int(counts[-1]),
           "bit_exact_vs_oracle": r1["bit_exact_vs_oracle"] and rk["bit_exact_vs_oracle"]}
    ref = native.ref("sciddicat_active")
    if ref is not None:
        t0 = time.perf_counter()
        native.ref_sciddicat(z, h, 0.5, 0.001, cpu_steps, True)
        cw = time.perf_counter() - t0
        out["cpu_baseline"] = cpu_record(cw / cpu_steps * 1e3, "ms/step", "reference",
                                         f"reference OpenCAL-OMP active-cells sciddicaT {n}x{n}, first {cpu_steps} "
                                         f"steps incl. init, {cw:.1f} s")
    return out


def bench_wet(n=4096, steps=100, seed=2024):
    """SciddicaT 4096^2 full grid with debris on every cell (h ~ U(0.5, 5) m,
    the headline's DEM): the minimisation runs everywhere, the tile mode
    declines and the engine takes its strip sweeps (one or two steps per
    launch, whichever its first-use timing of the grid shape found faster;
    one for this wet grid)."""
    import native
    from opencal_amd import synth
    from opencal_amd.models import SciddicaT
    z, _, _ = synth.sciddicat_inputs(n, n, seed)
    h = 0.5 + 4.5 * synth.uniform(seed + 1, n * n).reshape(n, n)
    m = SciddicaT(z, h, 0.5, 0.001, steps + 8, active=False)
    m.init()
    m.steps(1)
    m.run_steps(4)          # the sweep's shape is chosen on first use
    m.sync()
    m.timing(True)
    t0 = time.perf_counter()
    m.run_steps(steps)
    m.sync()
    wall = time.perf_counter() - t0
    ktime_ms, launches = m.kernel_time()
    m.timing(False)
    _, gh, _ = m.read()
    m.close()
    _, oh = native.oracle_sciddicat(z, h, 0.5, 0.001, steps + 5, False)
    wet = float((oh > 0.001).mean())
    return {"workload": f"sciddicaT_plain_wet_{n}x{n}_x{steps}", "gcups": n * n * steps / wall / 1e9,
            "ms_per_step": wall / steps * 1e3, "launches_per_step": launches / steps,
            "kernel_us_per_launch": ktime_ms / max(launches, 1) * 1e3,
            "achieved_gbs_per_launch": MIN_BYTES_PER_CELL_SCIDDICAT * n * n / (ktime_ms / max(launches, 1) * 1e-3) / 1e9,
            "wet_fraction_at_end": wet, "bit_exact_vs_oracle": bool(np.array_equal(gh, oh)),
            "note": "wall time with per-launch HIP events; achieved = 16 B per cell over the average launch "
                    "(a strip sweep reads z and h and writes h: 24 B per cell and launch)"}


def bench_life(n=512, steps=1000, seed=1):
    import native
    from opencal_amd import synth
    from opencal_amd.models import Life
    a0 = synth.life_inputs(n, n, seed)
    res = {}
    warm = Life(a0, 40)      # loads the kernels and sizes the runtime's pools outside the timed runs
    warm.run()
    warm.read()
    warm.close()
    for mode in ("run", "steps"):
        m = Life(a0, steps)
        t0 = time.perf_counter()
        if mode == "run":
            m.run()             # calRun2D: no callbacks between steps -> temporally blocked launches
        else:
            m.steps(steps)      # one calRunCAStep2D (one launch) per step
        got = m.read()
        wall = time.perf_counter() - t0
        m.close()
        res[mode] = (wall, got)
    want = native.oracle_life(a0, steps)
    alive = int(want.sum())
    cpu = None
    if native.ref("life") is not None:
        t0 = time.perf_counter()
        native.ref_life(a0, steps)
        cw = time.perf_counter() - t0
        cpu = cpu_record(n * n * steps / cw / 1e9, "GCUPS", "reference",
                         f"reference OpenCAL-OMP life (int substate) {n}x{n}, {steps} steps, {cw:.2f} s")
    return {"workload": f"life_{n}x{n}_x{steps}_uint8", "cpu_baseline": cpu,
            "calRun2D_us_per_step": res["run"][0] / steps * 1e6,
            "calRun2D_gcups": n * n * steps / res["run"][0] / 1e9,
            "per_step_launch_us_per_step": res["steps"][0] / steps * 1e6,
            "alive": alive, "bit_exact_vs_oracle": bool(np.array_equal(res["run"][1], want) and
                                                        np.array_equal(res["steps"][1], want)),
            "note": "wall time incl. upload/download of the 256 KiB board"}


def bench_batch(n_inst=256, steps=2000, seed=4, sequential=True):
    from opencal_amd import synth
    from opencal_amd.models import SciddicaT, sciddicat_batch
    specs = synth.batch_instances(n_inst, seed)
    insts = []
    for rows, cols, pr, pe, s in specs:
        z, h, _ = synth.sciddicat_inputs(rows, cols, s)
        insts.append((z, h, pr, pe))
    cells = sum(r * c for r, c, *_ in specs)
    log(f"batch: {n_inst} instances, {cells / 1e6:.1f} M cells")
    outs, ms = sciddicat_batch(insts, steps)
    if not sequential:
        return {"workload": f"sciddicaT_batch_{n_inst}_instances_128..1024_x{steps}", "cells": cells,
                "batched_gcups": cells * steps / (ms * 1e-3) / 1e9, "batched_s": ms * 1e-3}
    # the same instances one after another, each through calRun2D's step loop
    # (calhipRunSteps2D, the headline's driver)
    seq_s = 0.0
    exact = True
    for (z, h, pr, pe), got in zip(insts, outs):
        m = SciddicaT(z, h, pr, pe, steps + 2, False)
        m.init()
        m.steps(1)
        m.run_steps(2)   # upload, and the per-shape choice of the sweep, outside the timed steps
        m.sync()
        t0 = time.perf_counter()
        m.run_steps(steps - 3)
        m.sync()
        seq_s += (time.perf_counter() - t0) * steps / (steps - 3)
        _, hh, _ = m.read()
        m.close()
        exact = exact and bool(np.array_equal(hh, got))
    bat_gcups = cells * steps / (ms * 1e-3) / 1e9
    seq_gcups = cells * steps / seq_s / 1e9
    # CPU: the reference build on a sample of the instances, 100 steps each
    cpu = None
    import native
    if native.ref("sciddicat") is not None:
        sample = insts[:: max(1, len(insts) // 8)][:8]
        sc = sum(z.size for z, *_ in sample)
        t0 = time.perf_counter()
        for z, h, pr, pe in sample:
            native.ref_sciddicat(z, h, pr, pe, 100)
        cw = time.perf_counter() - t0
        cpu = cpu_record(sc * 100 / cw / 1e9, "GCUPS", "reference",
                         f"reference OpenCAL-OMP sciddicaT, {len(sample)} of the instances ({sc / 1e6:.1f} M cells) "
                         f"one after another, 100 steps each incl. init, {cw:.1f} s")
    return {"workload": f"sciddicaT_batch_{n_inst}_instances_128..1024_x{steps}", "cells": cells,
            "batched_gcups": bat_gcups, "sequential_gcups": seq_gcups, "speedup": bat_gcups / seq_gcups,
            "sequential_driver": "each instance through calhipRunSteps2D (calRun2D's loop, the headline's driver)",
            "batched_s": ms * 1e-3, "sequential_s": seq_s, "each_instance_equals_its_single_run": exact,
            "cpu_baseline": cpu}


def bench_adjoint(n=1024, steps=2000, ck=45, seed=5):
    import torch
    from opencal_amd import synth
    from opencal_amd.adjoint import forward_only, sciddicat_run
    z_np, h_np, (r0, c0, p) = synth.sciddicat_inputs(n, n, seed)
    dev = "cuda"
    z_dem = torch.tensor(z_np, dtype=torch.float64, device=dev)
    patch0 = torch.tensor(h_np[r0:r0 + p, c0:c0 + p].copy(), dtype=torch.float64, device=dev)
    target = forward_only(z_dem, patch0, (r0, c0), 0.6, 0.003, steps)      # hidden parameters
    forward_only(z_dem, patch0, (r0, c0), 0.5, 0.001, 10)                  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    forward_only(z_dem, patch0, (r0, c0), 0.5, 0.001, steps)
    torch.cuda.synchronize()
    fwd_s = time.perf_counter() - t0
    # the best plain run of the same forward: the C API model through
    # calRun2D's loop (the headline's driver)
    from opencal_amd.models import SciddicaT
    h_src = np.zeros_like(z_np)
    h_src[r0:r0 + p, c0:c0 + p] = h_np[r0:r0 + p, c0:c0 + p]
    m = SciddicaT(z_np, h_src, 0.5, 0.001, steps + 2, False)
    m.init()
    m.steps(1)
    m.run_steps(2)   # upload and the per-shape sweep choice outside the timed steps
    m.sync()
    t0 = time.perf_counter()
    m.run_steps(steps - 3)
    m.sync()
    plain_s = (time.perf_counter() - t0) * steps / (steps - 3)
    m.close()
    def grad_run():
        pr = torch.tensor(0.5, dtype=torch.float64, requires_grad=True)
        pe = torch.tensor(0.001, dtype=torch.float64, requires_grad=True)
        patch = patch0.clone().requires_grad_(True)
        hT = sciddicat_run(pr, pe, patch, z_dem, (r0, c0), steps, ck)
        loss = ((hT - target) ** 2).sum()
        loss.backward()
        return pr, pe, patch, loss
    grad_run()                        # warm-up: work buffers and checkpoint storage allocated once
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    pr, pe, patch, loss = grad_run()
    torch.cuda.synchronize()
    grad_s = time.perf_counter() - t0

    def L(prv):
        with torch.no_grad():
            return ((forward_only(z_dem, patch0, (r0, c0), prv, 0.001, steps) - target) ** 2).sum().item()
    h = 1e-6
    fd = (L(0.5 + h) - L(0.5 - h)) / (2 * h)
    g = pr.grad.item()
    return {"workload": f"sciddicaT_adjoint_{n}x{n}_x{steps}_ck{ck}", "forward_s": fwd_s,
            "gradient_s": grad_s, "gradient_over_forward": grad_s / fwd_s,
            "plain_calRun2D_forward_s": plain_s, "gradient_over_plain_forward": grad_s / plain_s,
            "loss": loss.item(),
            "dL_dpr": g, "dL_dpeps": pe.grad.item(), "patch_grad_norm": patch.grad.norm().item(),
            "fd_dL_dpr": fd, "fd_rel_err": abs(g - fd) / max(abs(fd), 1e-30)}


def bench_sciara(n=2048, seed=3, cpu_steps=200, cpu=True):
    """SCIARA-fv2 lava flow (config: 2048^2, Moore, active cells, 3 vents, 10
    emission intervals of 1000 steps; the model's stop condition ends the run
    after 10000 steps), stepped through calRunCAStep2D with the steering twin,
    model resident on the GPU after the first step."""
    import native
    from opencal_amd import synth
    from opencal_amd.models import Sciara
    inp = synth.sciara_inputs(n, n, seed)
    total = int(inp["emission_time"] / inp["params"][0]) * inp["rates"].shape[1]
    m = Sciara(inp, total + 10, active=True)
    m.init()
    m.steps(1)
    m.sync()
    t0 = time.perf_counter()
    ran = m.steps(total + 10)
    m.sync()
    wall = time.perf_counter() - t0
    res = m.read()
    st = m.stats()
    m.close()
    steps = ran + 1
    out = {"workload": f"sciara_fv2_{n}x{n}_active_x{steps}", "steps": steps, "ms_per_step": wall / ran * 1e3,
           "final_active_cells": res["active"],
           "lava_cells": int((res["h"] > 0).sum()), "solidified_cells": int((res["Msl"] > 0).sum()),
           "launches_per_step": st["launches"] / max(st["steps"], 1)}
    # lava volume balance (no lava reaches the border mask in this run)
    pac = inp["cell_size"] ** 2
    em = float((inp["rates"] * inp["params"][0] / pac).sum() * (inp["emission_time"] / inp["params"][0]))
    out["volume_balance_rel"] = float((res["h"].sum() + res["Msl"].sum() - em) / em)
    # CPU: the reference build (Sciara.cpp on the OpenCAL-OMP engine) on the
    # first cpu_steps steps of the same run, else the oracle restatement; the
    # cost per step is dominated by the full-grid active-set rebuilds and
    # barely depends on the lava extent
    # (the reference build runs in a subprocess: its OpenMP run has ended in
    # a segmentation fault on some hosts, which then costs only this leg)
    if not cpu:
        return out
    kind = "reference" if native.ref("sciara") is not None else "port"
    cw, note = None, ""
    if kind == "reference":
        cw, err = ref_leg_child(["sciara", str(n), str(seed), str(cpu_steps)])
        if cw is None:
            kind, note = "port", f"; the reference build's run failed ({err}), the oracle was timed instead"
    if cw is None:
        t0 = time.perf_counter()
        native.oracle_sciara(inp, cpu_steps, True, crmath=False)
        cw = time.perf_counter() - t0
    out["cpu_baseline"] = cpu_record(cw / cpu_steps * 1e3, "ms/step", kind,
                                     f"{'reference OpenCAL-OMP Sciara.cpp' if kind == 'reference' else 'oracle'} "
                                     f"{n}x{n} active cells, first {cpu_steps} steps incl. init, {cw:.1f} s"
                                     f"{note}")
    if kind == "port":
        out["cpu_baseline"]["cores"] = 1
    return out


def ref_leg(name, n, seed, steps):
    """The CPU leg of a workload on the reference build alone (no GPU use);
    returns the seconds it took."""
    import native
    from opencal_amd import synth
    if name != "sciara":
        raise ValueError(name)
    inp = synth.sciara_inputs(n, n, seed)
    t0 = time.perf_counter()
    native.ref_sciara(inp, steps, True)
    return time.perf_counter() - t0


def ref_leg_child(args, timeout_s=240):
    """ref_leg in a child process (`bench.py --ref-leg NAME N SEED STEPS`),
    waited for: (seconds, None), or (None, what went wrong)."""
    cmd = [sys.executable, "-X", "faulthandler", os.path.abspath(__file__), "--ref-leg", *args]
    try:
        p = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout_s)
    except subprocess.TimeoutExpired:   # run() kills and reaps the child
        return None, f"time-out after {timeout_s} s"
    if p.returncode == 0:
        for line in reversed(p.stdout.splitlines()):
            if line.startswith("{"):
                return float(json.loads(line)["seconds"]), None
    return None, f"exit status {p.returncode}"


def live_children():
    """Processes this run started that are still alive (each child of this
    script is waited for; a non-empty list names what a library left)."""
    try:
        import psutil
    except ImportError:
        return None
    out = []
    for p in psutil.Process().children(recursive=True):
        try:
            out.append({"pid": p.pid, "name": p.name(), "cmdline": " ".join(p.cmdline())[:200]})
        except psutil.Error:
            pass
    return out


def run_workload_child(name, timeout_s=300):
    """One secondary workload in a child process (`bench.py --only name`),
    waited for before returning: a crash or a time-out there becomes an error
    record in this line instead of ending the run, and the child's last
    stderr lines (with a Python traceback on a crash) come with it."""
    cmd = [sys.executable, "-X", "faulthandler", os.path.abspath(__file__), "--only", name]
    try:
        p = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout_s)
    except subprocess.TimeoutExpired as e:   # run() kills and reaps the child
        tail = (e.stderr or b"")
        tail = tail.decode(errors="replace") if isinstance(tail, bytes) else tail
        return {"error": f"time-out after {timeout_s} s", "stderr_tail": tail.splitlines()[-20:]}
    for line in reversed(p.stdout.splitlines()):
        if line.startswith("{"):
            try:
                rec = json.loads(line).get(name)
            except ValueError:
                break
            if p.returncode == 0 and rec is not None:
                return rec
            break
    return {"error": f"exit status {p.returncode}", "stderr_tail": p.stderr.splitlines()[-20:]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--size", type=int, default=4096)
    ap.add_argument("--cpu-steps", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--quick", action="store_true", help="headline only")
    ap.add_argument("--only", choices=["active", "wet", "life", "batch", "adjoint", "sciara"],
                    help="run one secondary workload alone and print its result (for profiling)")
    ap.add_argument("--batched-only", action="store_true",
                    help="with --only batch: the batched run alone, without the sequential and CPU legs")
    ap.add_argument("--ref-leg", nargs=4, metavar=("NAME", "N", "SEED", "STEPS"),
                    help="time a workload's CPU leg on the reference build alone (used as a child process)")
    args = ap.parse_args()
    if args.ref_leg:
        name, n, seed, steps = args.ref_leg
        print(json.dumps({"seconds": ref_leg(name, int(n), int(seed), int(steps))}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the engine runs on one MI355X; "
              "how the path would shard is a design note)", file=sys.stderr)
        sys.exit(2)

    import torch
    assert torch.cuda.is_available(), "needs a GPU"
    workloads = {"active": bench_active, "wet": bench_wet, "life": bench_life, "batch": bench_batch,
                 "adjoint": bench_adjoint, "sciara": bench_sciara}
    if args.only:
        fn = workloads[args.only]
        if args.only == "sciara" and args.no_cpu:
            fn = lambda: bench_sciara(cpu=False)   # noqa: E731
        if args.only == "batch" and args.batched_only:
            fn = lambda: bench_batch(sequential=False)   # noqa: E731
        print(json.dumps({args.only: fn()}), flush=True)
        return
    log(f"sciddicaT plain {args.size}^2: warmup {args.warmup}, timed {args.steps}")
    r = bench_sciddicat_plain(args.size, args.steps, args.warmup, batched=True)
    log(json.dumps(r))
    r1 = bench_sciddicat_plain(args.size, args.steps, args.warmup, batched=False)
    log(json.dumps(r1))
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
                   "driver": "calRun2D's step loop (calhipRunSteps2D), steering = the reference flow reset "
                             "with its device twin: up to four steps per launch"},
        "roofline": tiles_roofline(r, args.size),
        "per_step_driver": {
            "driver": "calRunCAStep2D per step (the engine groups four steps to a launch)",
            "gcups": round(r1["gcups"], 3),
            "ms_per_step": round(r1["wall_s"] / args.steps * 1e3, 4),
            "roofline": tiles_roofline(r1, args.size)},
    }
    log("pcie-inclusive run ...")
    out["pcie_inclusive"] = pcie_inclusive_sciddicat(args.size, args.steps)
    if not args.no_cpu:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline_sciddicat(args.size, args.cpu_steps)
    if not args.quick:
        wl = {}
        for name in workloads:
            log(f"{name} ...")
            t0 = time.perf_counter()
            wl[name] = run_workload_child(name)
            log(f"{name}: {json.dumps(wl[name])} ({time.perf_counter() - t0:.1f} s)")
        out["workloads"] = wl
    out["children_at_exit"] = live_children()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
