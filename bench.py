#!/usr/bin/env python3
"""Benchmark: particle-steps/s of the dusty-gas SPH step on the 2 x 2,812,500-particle tube,
with the cfg3 sweep and the cfg4 serving stream in the same JSON line.

Workload (cfg2): isothermal dusty shock tube, x in [-0.5,0.5], y,z in [0,0.2) periodic,
lattice spacing 0.002 / 0.004, density 1 / 0.125, dust-to-gas 1, K = 100, seeded jitter,
CFL 0.3 steps to t = 0.2, mixed precision (FP32 pair terms accumulated in FP64, an FP64 band
next to the frozen walls; FP64 drag, positions and conservation sums).  Inputs are resident
in HBM before the timed region.

A "step" is one pass of the whole per-step pipeline (sort, density, force, IDIC drag +
kick-drift, per-tube reductions) over both species.  The K timed steps are spread evenly over
the run to t = 0.2 (h and the band width change along it), each bracketed by device
synchronisations; every stage is bracketed by HIP events on the step stream, and the dominant
stage's average duration gives the roofline figure with the algorithmic bytes of its pass.
The whole run, graph-replayed, and the same run in FP64 are reported beside it.
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

HBM_PEAK_GBS = 8000.0

# Algorithmic bytes per particle of each pass (compulsory traffic, cell-sorted SoA):
# (gas, dust).  The idic stage fuses the cell reduction (60 / 52) with the drag +
# kick-drift + energy + conservation pass (108 / 96).  Per averaging cell of one neighbour
# cell it adds the cell start of each species (2 x 4 B) and a 40 B conservation / CFL
# partial per block of 32 cells (1.25 B): 9 B.
ALGO_BYTES = {
    "sort": (144, 136),
    "density": (28, 28),
    "force": (68, 0),
    "idic": (168, 148),
}
AVG_CELL_BYTES = 9
DEFAULT_SHAPE = (192, 1)   # the library's default block shape: piece size, waves per block
SORT_STAGES = ("key", "scan", "scatter", "rank_permute", "radius")


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


# kernels behind each pass, as named in the rocprofv3 summaries under profiles/
def pass_kernels(stage, shape=DEFAULT_SHAPE):
    """Kernel names (as rocprofv3 reports them) of one pass of the step."""
    tile, waves = shape
    return {
        # mixed precision with the FP64 wall band: one fused launch per pair pass
        "density": [f"k_pairs_fused<0, false, {tile}, {waves}>"],
        "force": [f"k_pairs_fused<1, false, {tile}, {waves}>"],
        "idic": ["k_idic_runs<1>"],
        "sort": ["k_key", "k_scan_reduce", "k_scan_top", "k_scan_down", "k_scatter", "k_rank_permute",
                 "k_window_max", "k_radius", "k_band", "k_perm_count", "k_perm_write"],
    }.get(stage, [])


def measured_traffic(workload, stage, shape=DEFAULT_SHAPE):
    """HBM bytes per launch of a pass from the committed FETCH_SIZE/WRITE_SIZE summary (or None)."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_{workload}.json")
    if not os.path.exists(path):
        return None
    kern = json.load(open(path))["kernels"]
    names = pass_kernels(stage, shape)
    if not names or any(n not in kern for n in names):
        return None
    return sum(kern[n]["read_bytes"] + kern[n]["write_bytes"] for n in names)


def measured_valu(workload, stage, shape=DEFAULT_SHAPE):
    """VALU issue fraction of a pass's kernel from the committed SQ-counter summary (or None)."""
    path = os.path.join(ROOT, "profiles", f"valu_{workload}.json")
    if not os.path.exists(path):
        return None
    kern = json.load(open(path))["kernels"]
    names = pass_kernels(stage, shape)
    if len(names) != 1 or names[0] not in kern:
        return None
    return kern[names[0]]["valu_issue_frac"]


def _oracle_rate(spec, threads, steps):
    import oracle

    oracle.set_threads(threads)
    o = oracle.OracleSim.from_spec(spec)
    o.step(0)  # prepare outside the timed region
    t0 = time.perf_counter()
    o.step(steps)
    dt = time.perf_counter() - t0
    n = sum(o.counts.values())
    return n * steps / dt, n, dt, oracle.get_threads()


def _cpu_model():
    """The host CPU's model name (lscpu, else /proc/cpuinfo)."""
    import subprocess

    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_threads():
    """The host threads this process may use: OMP_NUM_THREADS when set (the share of the host
    that goes with one GPU), else the cores the process may run on."""
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_baseline(spec, threads, steps=3):
    """The FP64 OpenMP oracle on the host cores: `steps` steps of the same tube from its set-up
    (threads bound one per core, OMP_PROC_BIND=close, OMP_PLACES=cores); and on one thread, one
    step of the cfg1-sized tube (a quarter of the particles)."""
    from dustyshock3d_amd import configs

    rate, n, dt, cores = _oracle_rate(spec, threads, steps)
    small = configs.cfg1()
    rate1, n1, dt1, _ = _oracle_rate(small, 1, 1)
    return {"value": rate, "unit": "particle-steps/s", "cores": cores, "kind": "port",
            "sample": f"{steps} steps of the full {n}-particle tube from its set-up, FP64 OpenMP oracle, "
                      f"{dt:.1f} s",
            "binding": {k: os.environ.get(k) for k in ("OMP_PROC_BIND", "OMP_PLACES", "OMP_NUM_THREADS")},
            "one_thread": {"value": rate1, "sample": f"1 step of the {n1}-particle cfg1 tube, {dt1:.1f} s"},
            "cpu_model": _cpu_model(), "host_cpus": os.cpu_count()}


def whole_run(spec, precision=None):
    """The tube from its set-up to t_end, graph-replayed, untimed by stage: (steps, seconds)."""
    from dustyshock3d_amd import Simulation

    if precision is not None:
        spec.precision = precision
    s = Simulation.from_spec(spec)
    s.step(0)
    t0 = time.perf_counter()
    s.step(1 << 20)
    dt = time.perf_counter() - t0
    steps = s.status()["steps"]
    s.close()
    return steps, dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="cfg2", choices=["cfg0", "cfg1", "cfg2"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--variant", type=int, default=1, help="pair kernels: 1 tiled (LDS), 0 direct")
    ap.add_argument("--no-autotune", action="store_true", help=f"keep the default block shape (piece size, waves) {DEFAULT_SHAPE}")
    ap.add_argument("--no-sweep", action="store_true", help="leave out the cfg3 sweep object")
    ap.add_argument("--no-serve", action="store_true", help="leave out the cfg4 serving object")
    ap.add_argument("--no-whole-run", action="store_true", help="leave out the whole-run figures")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    # the oracle's OpenMP threads (CPU baseline): one per core, close together
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")

    from dustyshock3d_amd import Simulation, abi, configs

    spec = getattr(configs, args.config)()
    t_setup = time.perf_counter()
    sim = Simulation.from_spec(spec)
    sim.set_pair_variant(args.variant)
    n_gas, n_dust = sim.counts[0], sim.counts[1]
    tune = None
    if args.variant == 1 and not args.no_autotune:
        # untimed: each block shape (piece size x waves per block) runs 1 + 2 ordinary steps of
        # the tube, the fastest is kept
        tune = sim.autotune(steps_per_trial=2)
    shape = (sim.block_shape()["tile"], sim.block_shape()["waves"])
    sim.step(args.warmup)
    setup_s = time.perf_counter() - t_setup
    st0 = sim.status()
    # the K timed steps, one every `stride` steps of the rest of the run; sim.step returns after
    # the device has finished, so each timed step is bracketed by synchronisations
    end_t = spec.t_end if spec.t_end < 1e29 else None
    left = (end_t - st0["t"]) / st0["dt"] if end_t else float(args.steps)
    stride = max(1, int(0.9 * left / max(args.steps, 1)))
    wall, stage_sum, steps_done, t_samples = 0.0, {}, 0, []
    for _ in range(args.steps):
        if stride > 1:
            sim.set_timing(False)
            sim.step(stride - 1)
        if sim.status()["done"]:
            break
        sim.set_timing(True)
        t_samples.append(sim.status()["t"])
        t0 = time.perf_counter()
        sim.step(1)
        wall += time.perf_counter() - t0
        steps_done += 1
        for k, v in sim.kernel_times().items():
            stage_sum[k] = stage_sum.get(k, 0.0) + v
    sim.set_timing(False)
    st1 = sim.status()
    kt = {k: v / max(steps_done, 1) for k, v in stage_sum.items()}
    if steps_done != args.steps:
        print(f"warning: only {steps_done} of {args.steps} steps ran (tube finished)", file=sys.stderr)
    n_avg = sim.grid()["n_avg"]
    sim.close()
    value = (n_gas + n_dust) * steps_done / wall

    groups = {"sort": sum(kt.get(k, 0.0) for k in SORT_STAGES)}
    for k in ("density", "force", "idic", "finalize"):
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

    runs = None
    if not args.no_whole_run and end_t:
        n_mix, s_mix = whole_run(getattr(configs, args.config)())
        n_64, s_64 = whole_run(getattr(configs, args.config)(), abi.FP64)
        runs = {"mixed": {"steps": n_mix, "seconds": round(s_mix, 3), "ms_per_step": round(s_mix * 1e3 / n_mix, 4),
                          "particle_steps_per_s": (n_gas + n_dust) * n_mix / s_mix},
                "fp64": {"steps": n_64, "seconds": round(s_64, 3), "ms_per_step": round(s_64 * 1e3 / n_64, 4),
                         "particle_steps_per_s": (n_gas + n_dust) * n_64 / s_64},
                "note": "set-up to t = 0.2 as one graph-replayed run, no stage events"}

    cpu = None
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or cpu_threads()
        cpu = cpu_baseline(spec, threads, args.cpu_steps)

    sweep = None if args.no_sweep else sweep_result(SweepArgs())
    serve = None if args.no_serve else serve_result(ServeArgs())

    out = {
        "metric": "particle-steps/s at 2x2.8M (HBM roofline %); sweep p50/p99 latency vs serial",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": steps_done,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / max(steps_done, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 pair terms accumulated in f64 + f64 wall band, drag, positions, reductions"
                 if spec.precision == 1 else "f64",
        "data": "synthetic (seeded lattice shock tube)",
        "config": {"workload": f"{args.config}: dusty shock tube, {n_gas} gas + {n_dust} dust particles, "
                               f"y,z width {spec.y_width}, dx {spec.dx_left}/{2 * spec.dx_left}, K={spec.K}, CFL 0.3"
                               + (f", to t = {end_t}" if end_t else ""),
                   "particles": n_gas + n_dust, "averaging_cells": n_avg,
                   "timed_steps": f"{steps_done} single steps, one every {stride} steps of the run, "
                                  f"t = {t_samples[0]:.4f} to {t_samples[-1]:.4f}" if t_samples else None,
                   "t_after": st1["t"], "setup_s": round(setup_s, 2),
                   "pair_kernels": "tiled" if args.variant == 1 else "direct",
                   "autotune": tune},
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(args.config, dominant, shape)
                     if args.variant == 1 else None,
                     "traffic_source": f"profiles/hbm_traffic_{args.config}.json (rocprofv3 FETCH_SIZE/WRITE_SIZE)",
                     "algo_bytes_per_launch": dom_bytes, "avg_ms": dom_ms,
                     # the pair passes are compute-bound: the share of VALU issue slots they use
                     "valu_issue_frac": measured_valu(args.config, dominant, shape)
                     if args.variant == 1 else None,
                     "valu_source": f"profiles/valu_{args.config}.json (rocprofv3 SQ_INSTS_VALU)"},
        "step_roofline": {"algo_bytes": step_bytes, "event_ms": step_ms,
                          "frac": step_bytes / (wall / max(steps_done, 1)) / 1e9 / HBM_PEAK_GBS},
        "stages_ms": {k: round(v, 4) for k, v in kt.items()},
        "passes": per_stage,
        "whole_run": runs,
        "cpu_baseline": cpu,
        "sweep": sweep,
        "serve": serve,
    }
    print(json.dumps(out))


def run_alone(spec, variant=1):
    """One tube alone on the GPU (the plain one-tube-at-a-time form): particle-steps, seconds."""
    from dustyshock3d_amd import Simulation

    s = Simulation.from_spec(spec)
    s.set_pair_variant(variant)
    s.step(0)
    t0 = time.perf_counter()
    while not s.status()["done"]:
        s.step(8)
    dt = time.perf_counter() - t0
    st = s.status()
    return sum(s.counts.values()) * st["steps"], dt


class SweepArgs:
    """Defaults of the cfg3 sweep as the main benchmark line runs it; `split` batches: the
    alternative policy reported beside it (the lighter half by estimated work in a batch of
    its own, beside the rest)."""
    variant, tubes, chunk, serial_stride, split = 1, 512, 16, 8, 2


class ServeArgs:
    """Defaults of the cfg4 stream as the main benchmark line runs it.  heavy: requests above
    this many estimated particle-steps start one at a time (the large tubes, ~3.9e8; the sweep
    tubes stay below 4e7)."""
    variant, requests, load, sat_rate, budget, epoch, seed, heavy = 1, 60, 0.8, 0.0, 200e9, 8, 7, 1e8


def _serial_model(sample):
    """Least-squares fit of one-tube-at-a-time run times, seconds = steps x (a + b x particles),
    to measured (particles, steps, seconds) samples (relative residuals); returns (a, b) and
    the largest relative deviation of the fit from a sample."""
    P = np.array([x[0] for x in sample], float)
    S = np.array([x[1] for x in sample], float)
    D = np.array([x[2] for x in sample], float)
    A = np.stack([S / D, S * P / D], axis=1)
    (a, b), *_ = np.linalg.lstsq(A, np.ones_like(D), rcond=None)
    dev = float(np.max(np.abs(S * (a + b * P) / D - 1.0)))
    return float(a), float(b), dev


def sweep_result(args) -> dict:
    """cfg3: 512 independent tubes, each run to t=0.2, all submitted at once: in one batch, and
    (reported beside it) lightest half first in a concurrent batch of its own, against the same
    tubes run one at a time on the same GPU.  Every `serial_stride`-th tube is run alone and
    timed; the serial queue's times to solution (one tube after another in submission order)
    come from a per-step cost fitted to those runs."""
    from dustyshock3d_amd import configs
    from dustyshock3d_amd.batch import run_sweep, sweep_groups
    from dustyshock3d_amd.sim import tube_count

    specs = configs.cfg3(args.tubes)
    parts = [2 * tube_count(sp) for sp in specs]
    runs = {}
    for nb in (1, args.split):
        t0 = time.perf_counter()
        r = run_sweep(specs, batches=nb, chunk=args.chunk, pair_variant=args.variant, concurrent=True)
        wall = time.perf_counter() - t0
        ps = sum(p * st for p, st in zip(parts, r["steps"]))
        tts = np.array(r["finish_s"])
        runs[nb] = {"value": ps / wall, "wall_s": wall, "particle_steps": ps, "steps": r["steps"],
                    "time_to_solution_s": {"p50": float(np.percentile(tts, 50)),
                                           "p99": float(np.percentile(tts, 99))}}
    steps = runs[1]["steps"]
    if runs[args.split]["steps"] != steps:
        raise RuntimeError("a tube took a different number of steps in a different batch")
    sample_ids = list(range(0, len(specs), args.serial_stride))
    sample, sps, sdt = [], 0.0, 0.0
    for i in sample_ids:
        p, d = run_alone(specs[i], args.variant)
        sps += p
        sdt += d
        sample.append((parts[i], p / parts[i], d))
    serial_rate = sps / sdt
    a, b, dev = _serial_model(sample)
    service = np.array([st * (a + b * p) for p, st in zip(parts, steps)])
    serial_tts = np.cumsum(service)
    head, alt = runs[1], runs[args.split]
    return {
        "metric": "sweep particle-steps/s (cfg3, 512 tubes in one batch)",
        "value": head["value"], "unit": "particle-steps/s", "n_gpus": 1, "higher_is_better": True,
        "vs_serial_gpu": head["value"] / serial_rate,
        "serial_gpu_particle_steps_per_s": serial_rate,
        "serial_sample": f"{len(sample_ids)} of {len(specs)} tubes run one at a time, {sdt:.1f} s",
        "wall_s": head["wall_s"], "particle_steps": head["particle_steps"],
        "tubes": len(specs), "steps_min_max": [min(steps), max(steps)],
        "time_to_solution_s": head["time_to_solution_s"],
        "serial_fifo_time_to_solution_s": {"p50": float(np.percentile(serial_tts, 50)),
                                           "p99": float(np.percentile(serial_tts, 99)),
                                           "model": f"seconds = steps x ({a:.3e} + {b:.3e} x particles), "
                                                    f"fitted to the {len(sample)} tubes run alone "
                                                    f"(largest deviation {dev:.1%})"},
        "lightest_half_first": {"batches": [len(g) for g in sweep_groups(specs, args.split)],
                                "value": alt["value"], "wall_s": alt["wall_s"],
                                "time_to_solution_s": alt["time_to_solution_s"],
                                "policy": "tubes in ascending estimated work (particles x CFL steps at rest); "
                                          "the lighter half in a batch of its own, run concurrently beside "
                                          "the rest (own stream and host thread)"},
        "dtype": "f32 pair terms accumulated in f64 + f64 wall band, drag, positions, reductions", "data": "synthetic",
        "config": {"workload": "cfg3 sweep: y,z in [0,0.05], dx log-uniform [0.0025,0.01], K log-uniform "
                               "[0.01,1000], dust-to-gas uniform [0.01,10], each to t=0.2, all submitted at t=0",
                   "pair_kernels": "tiled" if args.variant == 1 else "direct"},
    }


def sweep_mode(args):
    print(json.dumps(sweep_result(args)))


def serve_result(args) -> dict:
    """cfg4: Poisson arrivals at a fraction of the measured saturated throughput, 95% sweep tubes
    and 5% large tubes, continuous batching under an HBM budget; the same request stream replayed
    one request at a time (FIFO, measured service times) gives the serial comparison."""
    from dustyshock3d_amd import abi, configs
    from dustyshock3d_amd.batch import Server

    rng = np.random.default_rng(args.seed)
    reqs = []
    for i in range(args.requests):
        if rng.uniform() < 0.05:
            reqs.append(configs.large_tube(i, precision=abi.MIXED))
        else:
            reqs.append(configs.sweep_tube(10000 + i, precision=abi.MIXED))
    # service time of every request alone (serial plain form), measured
    service, psteps = [], []
    for sp in reqs:
        p, d = run_alone(sp, args.variant)
        service.append(d)
        psteps.append(p)
    service = np.array(service)
    psteps = np.array(psteps)
    sat_rate = args.sat_rate if args.sat_rate > 0 else psteps.sum() / service.sum()
    # arrivals at `load` of the saturated rate, in requests/s
    lam = args.load * sat_rate / psteps.mean()
    arrivals = np.cumsum(rng.exponential(1.0 / lam, size=len(reqs)))
    # serial FIFO replay with the measured service times
    t_free, serial_tts = 0.0, []
    for a, s in zip(arrivals, service):
        start = max(a, t_free)
        t_free = start + s
        serial_tts.append(t_free - a)
    serial_tts = np.array(serial_tts)
    # continuous batching
    t0 = time.perf_counter()
    clock = lambda: time.perf_counter() - t0  # noqa: E731
    sv = Server(hbm_budget=args.budget, steps_per_epoch=args.epoch, pair_variant=args.variant, clock=clock,
                heavy_work=args.heavy)
    ids, k = [], 0
    while True:
        now = clock()
        while k < len(reqs) and arrivals[k] <= now:
            ids.append(sv.submit(reqs[k], float(arrivals[k])))
            k += 1
        r, q = sv.pump()
        if k == len(reqs) and r == 0 and q == 0:
            break
        if r == 0 and q == 0 and k < len(reqs):
            time.sleep(max(0.0, min(0.01, arrivals[k] - clock())))
    wall = clock()
    polls = [sv.poll(i) for i in ids]
    done = [p for p in polls if p["state"] == abi.REQ_DONE]
    tts = np.array([p["finished"] - p["arrival"] for p in done])
    ps = sum(p["particle_steps"] for p in done)
    st = sv.stats()
    return {
        "metric": "serving particle-steps/s and time-to-solution (cfg4 stream, continuous batching)",
        "value": ps / wall, "unit": "particle-steps/s", "n_gpus": 1, "higher_is_better": True,
        "requests": len(reqs), "completed": len(done),
        "large_requests": int(sum(1 for sp in reqs if sp.y_width > 0.05)),
        "offered_load": args.load, "arrival_rate_per_s": lam,
        "saturated_serial_rate": sat_rate,
        "time_to_solution_s": {"p50": float(np.percentile(tts, 50)), "p99": float(np.percentile(tts, 99))},
        "serial_fifo_time_to_solution_s": {"p50": float(np.percentile(serial_tts, 50)),
                                           "p99": float(np.percentile(serial_tts, 99))},
        "serial_particle_steps_per_s": float(psteps.sum() / service.sum()),
        "wall_s": wall, "rebuilds": st["rebuilds"], "epochs": st["epochs"],
        "rebuild_s": st["rebuild_s"], "advance_s": st["advance_s"], "hbm_budget": args.budget,
        "dtype": "f32 pair terms accumulated in f64 + f64 wall band, drag, positions, reductions", "data": "synthetic",
        "config": {"workload": "cfg4: Poisson arrivals, 95% cfg3-distribution tubes, 5% large tubes "
                               "(y,z in [0,0.1], dx 0.002), each to t=0.2", "steps_per_epoch": args.epoch,
                   "heavy_work": args.heavy,
                   "policy": "continuous batching in arrival order under the HBM budget; requests above "
                             "heavy_work particle-steps start one at a time"},
    }


def serve_mode(args):
    print(json.dumps(serve_result(args)))


def copy_mode(args):
    """A plain device-to-device copy of known size: the reference for scaling the HBM
    counters (FETCH_SIZE / WRITE_SIZE) of the step kernels to bytes."""
    import torch

    n = int(args.copy_bytes) // 4
    a = torch.ones(n, device="cuda", dtype=torch.float32)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.copy_reps):
        b.copy_(a)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.copy_reps
    print(json.dumps({"metric": "device copy", "bytes_read": 4 * n, "bytes_written": 4 * n,
                      "reps": args.copy_reps, "GBs": 8 * n / dt / 1e9}))


if __name__ == "__main__":
    if "--mode" in sys.argv:
        ap = argparse.ArgumentParser()
        ap.add_argument("--mode", choices=["sweep", "serve", "copy"], required=True)
        ap.add_argument("--copy-bytes", type=float, default=1 << 30)
        ap.add_argument("--copy-reps", type=int, default=5)
        ap.add_argument("--gpus", type=int, default=1)
        ap.add_argument("--variant", type=int, default=1)
        ap.add_argument("--tubes", type=int, default=512)
        ap.add_argument("--chunk", type=int, default=16)
        ap.add_argument("--serial-stride", type=int, default=8)
        ap.add_argument("--split", type=int, default=2, help="sweep: batches of the lightest-first policy")
        ap.add_argument("--requests", type=int, default=60)
        ap.add_argument("--load", type=float, default=0.8)
        ap.add_argument("--sat-rate", type=float, default=0.0)
        ap.add_argument("--budget", type=float, default=200e9)
        ap.add_argument("--epoch", type=int, default=8)
        ap.add_argument("--seed", type=int, default=7)
        ap.add_argument("--heavy", type=float, default=1e8)
        a = ap.parse_args()
        import numpy as np  # noqa: F811
        {"sweep": sweep_mode, "serve": serve_mode, "copy": copy_mode}[a.mode](a)
    else:
        main()
