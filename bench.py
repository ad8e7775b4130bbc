"""Benchmark: SU(3) link updates/s and % of HBM roofline at 32^4.

Workload (one "step"): one iteration of quenched SU(3) ensemble generation on a
32^4 lattice at beta=6.0 from a cold start, FP32 links in 12-real storage --
1 heatbath sweep + 4 overrelaxation sweeps (8 (parity, mu) launches each),
reunitarisation, plaquette and Polyakov loop with FP64 accumulation -- captured in
a hipGraph and replayed.  value = link updates per second of the whole job
(5 * 4V updates per step).  Inputs are resident in HBM before timing starts.

Prints ONE JSON line.  Extra objects:
  roofline      dominant kernel (the overrelaxation launch): algorithmic bytes
                9 * 48 B per link update x V/2 updates per launch, divided by the
                launch's average duration measured with HIP events on its stream
  cpu_baseline  the FP64 oracle on this host's cores, one iteration at 32^4
  plain_fp64    the plain single-lattice path (FP64 18-real, eager launches)
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

PEAK_HBM_GBS = 8000.0
DIMS = (32, 32, 32, 32)
BETA = 6.0


def link_bytes(nc, storage, prec):
    return (2 * nc * nc if storage == 18 else storage) * prec


def time_launches(sun_amd, lat, stream, hb: bool, reps: int) -> float:
    """average duration (ms) of one (parity, mu) update launch, HIP events on its stream"""
    t = sun_amd.Timer()
    for _ in range(2):  # warm
        if hb:
            lat.heatbath_sub(1, 0, 0, 0, stream)
        else:
            lat.overrelax_sub(0, 0, stream)
    t.start(stream)
    for r in range(reps):
        parity, mu = (r // 4) % 2, r % 4
        if hb:
            lat.heatbath_sub(1, r, parity, mu, stream)
        else:
            lat.overrelax_sub(parity, mu, stream)
    t.stop(stream)
    ms = t.ms() / reps
    t.close()
    return ms


def cpu_baseline(min_seconds: float = 10.0, max_iters: int = 20):
    """FP64 oracle on this host's cores: whole iterations (1 HB + 4 OR sweeps) of the
    same 32^4 workload until at least min_seconds have passed."""
    from oracle import oracle as o
    cores = min(os.cpu_count() or 1, 16)
    o.set_threads(cores)
    U = o.cold_start(3, DIMS)
    # one heatbath sweep first so the timed iterations start from a non-trivial field
    o.heatbath_sweep(U, DIMS, BETA, 2, 0)
    n = 0
    t0 = time.perf_counter()
    while n < max_iters and (n == 0 or time.perf_counter() - t0 < min_seconds):
        o.heatbath_sweep(U, DIMS, BETA, 2, 1 + n)
        for _ in range(4):
            o.overrelax_sweep(U, DIMS)
        n += 1
    dt = time.perf_counter() - t0
    V = int(np.prod(DIMS))
    return {"value": n * 5 * 4 * V / dt, "unit": "link-updates/s", "cores": cores, "kind": "port",
            "sample": f"FP64 oracle (C, OpenMP, {cores} threads): {n} iterations of 1 heatbath + 4 "
                      f"overrelaxation sweeps of SU(3) 32^4 beta=6.0 (after one heatbath sweep from cold), "
                      f"{dt:.1f} s"}


def kernels_only(sun_amd):
    """For rocprofv3 counter passes: on the 32^4 FP32 12-real field after a few
    iterations, 2 copies of the field (calibration: a plain copy of known size),
    then 8 overrelaxation launches, 8 heatbath launches, reunitarize, plaquette,
    Polyakov loop."""
    ctx = sun_amd.Context(0, 0, 0)
    lat = sun_amd.Lattice(ctx, 3, DIMS, BETA, 12, 4)
    tmp = sun_amd.Lattice(ctx, 3, DIMS, BETA, 12, 4)
    s = sun_amd.new_stream()
    lat.cold_start(s)
    lat.generate(seed=2, n_iter=3, n_or=4, reunit_every=1, stream=s)
    for _ in range(2):
        tmp.copy_from(lat, s)
    for parity in (0, 1):
        for mu in range(4):
            lat.overrelax_sub(parity, mu, s)
    for parity in (0, 1):
        for mu in range(4):
            lat.heatbath_sub(2, 3, parity, mu, s)
    lat.reunitarize(s)
    lat.plaquette(s)
    lat.polyakov(s)
    sun_amd.sync(s)
    print(f"kernels-only: field {lat.bytes} bytes, copy calibration size {tmp.bytes} bytes")
    tmp.close()
    lat.close()
    ctx.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the cpu_baseline leg")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain FP64 GPU leg")
    ap.add_argument("--kernels-only", action="store_true",
                    help="fixed short launch sequence for profiler counter passes (no JSON line)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {args.gpus} is not supported "
              "(lattices/jobs shard across GPUs without communication; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import sun_amd
    if args.kernels_only:
        kernels_only(sun_amd)
        return
    ctx = sun_amd.Context(0, 0, 1)
    V = int(np.prod(DIMS))
    updates_per_step = 5 * 4 * V
    storage, prec = 12, 4
    lat = sun_amd.Lattice(ctx, 3, DIMS, BETA, storage, prec)
    stream = sun_amd.new_stream()
    lat.cold_start(stream)
    sun_amd.sync(stream)

    # warmup (tunes launch shapes, builds and replays the graph)
    lat.generate(seed=2, n_iter=args.warmup, n_or=4, reunit_every=1, first_sweep=0, stream=stream)
    sun_amd.sync(stream)
    t0 = time.perf_counter()
    obs = lat.generate(seed=2, n_iter=args.steps, n_or=4, reunit_every=1, first_sweep=args.warmup, stream=stream)
    sun_amd.sync(stream)
    wall = time.perf_counter() - t0
    value = updates_per_step * args.steps / wall

    # dominant kernel: per-launch durations on the same stream
    or_ms = time_launches(sun_amd, lat, stream, hb=False, reps=64)
    hb_ms = time_launches(sun_amd, lat, stream, hb=True, reps=32)
    B = link_bytes(3, storage, prec)
    per_launch = 9 * B * (V // 2)
    or_gbs = per_launch / (or_ms * 1e-3) / 1e9
    hb_gbs = per_launch / (hb_ms * 1e-3) / 1e9
    iter_bytes = 4 * V * B * (45 + 2 + 1 + 0.25)
    step_ms = wall / args.steps * 1e3

    out = {
        "metric": "SU(3) link updates/s + % HBM roofline at 32^4",
        "value": value,
        "unit": "link-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": step_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (cold start, Philox heatbath)",
        "config": {"workload": "SU(3) 32^4 beta=6.0 cold start, FP32 12-real links; step = 1 heatbath + "
                               "4 overrelaxation sweeps + reunitarize + plaquette + Polyakov (FP64 sums), "
                               "hipGraph replay",
                   "lattice": list(DIMS), "beta": BETA, "storage": storage, "precision": "f32",
                   "updates_per_step": updates_per_step},
        "roofline": {"bound": "hbm", "achieved": or_gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": or_gbs / PEAK_HBM_GBS, "traffic": None,
                     "kernel": "overrelaxation (parity, mu) launch",
                     "algorithmic_bytes_per_launch": per_launch, "avg_launch_ms": or_ms,
                     "heatbath_launch_ms": hb_ms, "heatbath_gbs": hb_gbs,
                     "iteration_frac": iter_bytes / (step_ms * 1e-3) / 1e9 / PEAK_HBM_GBS},
        "observables": {"plaquette_last": float(obs[-1, :2].mean()), "polyakov_last": [float(obs[-1, 2]),
                                                                                        float(obs[-1, 3])]},
        "device": ctx.device_name(),
    }
    lat.close()

    if not args.no_plain:
        with sun_amd.Lattice(ctx, 3, DIMS, BETA, 18, 8) as pl:
            pl.cold_start(stream)
            pl.generate(seed=2, n_iter=2, n_or=4, reunit_every=1, use_graph=False, stream=stream)
            n = 10
            t0 = time.perf_counter()
            pl.generate(seed=2, n_iter=n, n_or=4, reunit_every=1, first_sweep=2, use_graph=False, stream=stream)
            sun_amd.sync(stream)
            pv = updates_per_step * n / (time.perf_counter() - t0)
        out["plain_fp64"] = {"value": pv, "unit": "link-updates/s",
                             "config": "FP64 18-real, one kernel per (parity, mu), eager launches",
                             "gain": value / pv}
    if not args.no_cpu:
        cb = cpu_baseline()
        out["cpu_baseline"] = cb
        out["gain_vs_cpu"] = value / cb["value"]
    ctx.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
