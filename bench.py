#!/usr/bin/env python3
"""Benchmark: particle-steps/s of the dusty-gas SPH step on the 2 x 2,812,500-particle tube.

Workload (cfg2): isothermal dusty shock tube, x in [-0.5,0.5], y,z in [0,0.2) periodic,
lattice spacing 0.002 / 0.004, density 1 / 0.125, dust-to-gas 1, K = 100, seeded jitter,
CFL 0.3 step, mixed precision (FP32 pair sums with Kahan compensation; FP64 drag,
positions and conservation sums).  Inputs are resident in HBM before the timed region.

A "step" is one pass of the whole per-step pipeline (sort, density, force, IDIC drag +
kick-drift, per-tube reductions) over both species.  Every stage is bracketed by HIP
events on the step stream; the dominant stage's average duration gives the roofline
figure with the algorithmic bytes of its pass.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# Algorithmic bytes per particle of each pass (compulsory traffic, cell-sorted SoA):
# (gas, dust).  The idic stage fuses the cell reduction (60 / 52) with the drag +
# kick-drift + energy + conservation pass (108 / 96) and adds 352 B per averaging cell.
ALGO_BYTES = {
    "sort": (144, 136),
    "density": (28, 28),
    "force": (68, 0),
    "idic": (168, 148),
}
AVG_CELL_BYTES = 352
SORT_STAGES = ("begin", "key", "scan", "scatter", "rank_permute", "cellmax", "radius")


def stage_bytes(stage, n_gas, n_dust, n_avg):
    if stage == "sort":
        g, d = ALGO_BYTES["sort"]
        return g * n_gas + d * n_dust
    if stage in ALGO_BYTES:
        g, d = ALGO_BYTES[stage]
        b = g * n_gas + d * n_dust
        if stage == "idic":
            b += AVG_CELL_BYTES * n_avg
        return b
    return 0


def cpu_baseline(spec, threads, steps=1):
    """The FP64 OpenMP oracle on the host cores: `steps` steps of the same tube."""
    import oracle

    oracle.set_threads(threads)
    o = oracle.OracleSim.from_spec(spec)
    o.step(0)  # prepare outside the timed region
    t0 = time.perf_counter()
    o.step(steps)
    dt = time.perf_counter() - t0
    n = sum(o.counts.values())
    return {"value": n * steps / dt, "unit": "particle-steps/s", "cores": oracle.get_threads(),
            "kind": "port",
            "sample": f"{steps} step(s) of the full {n}-particle tube (first step from the initial state), "
                      f"FP64 OpenMP oracle, {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="cfg2", choices=["cfg0", "cfg1", "cfg2"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=1)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--variant", type=int, default=1, help="pair kernels: 1 tiled (LDS), 0 direct")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    from dustyshock3d_amd import Simulation, configs

    spec = getattr(configs, args.config)(max_steps=0)
    t_setup = time.perf_counter()
    sim = Simulation.from_spec(spec)
    sim.set_pair_variant(args.variant)
    n_gas, n_dust = sim.counts[0], sim.counts[1]
    sim.step(args.warmup)
    setup_s = time.perf_counter() - t_setup
    st0 = sim.status()
    sim.set_timing(True)
    t0 = time.perf_counter()
    sim.step(args.steps)
    wall = time.perf_counter() - t0
    st1 = sim.status()
    kt = sim.kernel_times()
    steps_done = st1["steps"] - st0["steps"]
    if steps_done != args.steps:
        print(f"warning: only {steps_done} of {args.steps} steps ran (tube finished)", file=sys.stderr)
    n_avg = sim.grid()["n_avg"]
    value = (n_gas + n_dust) * steps_done / wall

    groups = {"sort": sum(kt.get(k, 0.0) for k in SORT_STAGES)}
    for k in ("density", "force", "idic", "finalize", "flip"):
        groups[k] = kt.get(k, 0.0)
    dominant = max(("density", "force", "idic", "sort"), key=lambda k: groups[k])
    dom_ms = groups[dominant]
    dom_bytes = stage_bytes(dominant, n_gas, n_dust, n_avg)
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9 if dom_ms > 0 else 0.0
    step_ms = sum(kt.values())
    step_bytes = sum(stage_bytes(k, n_gas, n_dust, n_avg) for k in ("sort", "density", "force", "idic"))
    per_stage = {}
    for k in ("sort", "density", "force", "idic"):
        b = stage_bytes(k, n_gas, n_dust, n_avg)
        per_stage[k] = {"ms": round(groups[k], 4), "algo_GBs": round(b / (groups[k] * 1e-3) / 1e9, 1)
                        if groups[k] > 0 else None,
                        "hbm_frac": round(b / (groups[k] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) if groups[k] > 0 else None}

    cpu = None
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(spec, threads, args.cpu_steps)

    out = {
        "metric": "particle-steps/s at 2x2.8M (HBM roofline %)",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": steps_done,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / max(steps_done, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 pair sums (Kahan) + f64 drag/positions/reductions" if spec.precision == 1 else "f64",
        "data": "synthetic (seeded lattice shock tube)",
        "config": {"workload": f"{args.config}: dusty shock tube, {n_gas} gas + {n_dust} dust particles, "
                               f"y,z width {spec.y_width}, dx {spec.dx_left}/{2 * spec.dx_left}, K={spec.K}, CFL 0.3",
                   "particles": n_gas + n_dust, "averaging_cells": n_avg, "t_at_start": st0["t"],
                   "dt": st1["dt"], "setup_s": round(setup_s, 2),
                   "pair_kernels": "tiled" if args.variant == 1 else "direct"},
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "algo_bytes_per_launch": dom_bytes, "avg_ms": dom_ms},
        "step_roofline": {"algo_bytes": step_bytes, "event_ms": step_ms,
                          "frac": step_bytes / (wall / max(steps_done, 1)) / 1e9 / HBM_PEAK_GBS},
        "stages_ms": {k: round(v, 4) for k, v in kt.items()},
        "passes": per_stage,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
