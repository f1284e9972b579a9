#!/usr/bin/env python3
"""Benchmark of the virtual-pipe water step on one MI355X.

Workload (the headline metric's configuration): one 4096 x 4096 FP32 grid,
seeded 6-octave value-noise fBm terrain scaled to 0-100 m plus a uniform
0.5 m water sheet, closed walls, g = 9.81, dt = 0.02 s, cell 1 m, pipe area
1 m^2, damping 0.995.  A bench "step" is one simulation step of the whole
grid; --steps K steps are timed (default 1000), after --warmup W untimed ones.

Printed: one JSON line with cell-updates/s, the roofline of the dominant
kernel (algorithmic bytes per launch / mean launch time from HIP events on the
simulation's stream) and a CPU baseline (the FP32 restatement under oracle/
on the host cores, on a bounded sample of the same grid).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_BPS = 8.0e12          # MI355X HBM3E, published
BYTES_PER_CELL_UPDATE = 44     # read b, d, 4 f (24 B) + write d, 4 f (20 B), FP32
PARAMS = dict(dt=0.02, gravity=9.81, cell_size=1.0, pipe_area=1.0, damping=0.995)
TRAFFIC_FILE = ROOT / "profiles" / "traffic.json"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(b, d, params, target_s=12.0):
    """FP32 restatement on all host cores, on a bounded number of steps of
    the same grid (step count calibrated from two warm steps to about
    target_s seconds)."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import tws_oracle
    threads = tws_oracle.set_threads(min(os.cpu_count() or 1, 16))
    F, K, Cc = params
    dd, ff, _ = tws_oracle.step(b, d, None, F, K, Cc, n=1)  # warm (threads, pages)
    t0 = time.perf_counter()
    dd, ff, _ = tws_oracle.step(b, dd, ff, F, K, Cc, n=2)
    one = (time.perf_counter() - t0) / 2
    steps = int(max(1, min(1000, target_s / max(one, 1e-6))))
    t0 = time.perf_counter()
    tws_oracle.step(b, dd, ff, F, K, Cc, n=steps)
    el = time.perf_counter() - t0
    H, W = b.shape
    return {
        "value": W * H * steps / el,
        "unit": "cell-updates/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{W}x{H} grid, {steps} steps after 3 warm steps ({el:.1f} s), "
                  "oracle/tws_oracle.c FP32 restatement, OpenMP over rows",
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--grid", type=int, default=4096)
    ap.add_argument("--kernel", type=int, default=0, help="TWS_KERNEL_* (0 = auto)")
    ap.add_argument("--k", type=int, default=0, help="steps per launch (0 = auto)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run the 1 GiB STREAM copy kernel (reference for PMC byte counts)")
    ap.add_argument("--extra", default="small,batched,served,large",
                    help="comma list of extra workloads in the same line: small, batched, "
                         "served, large ('' for none)")
    ap.add_argument("--served-requests", type=int, default=200)
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} is not supported; this build measures one "
              "GPU (the step of one grid runs on one device). Use --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import terrainwatersim_amd as tws

    if args.copy_calibration:
        gbps = tws.stream_copy_gbps(1 << 30, 10)
        print(json.dumps({"stream_copy_bytes": 1 << 30, "iterations": 11, "GBps": gbps}))
        return

    N = args.grid
    log(f"bench: generating {N}x{N} fBm terrain (seed 2)")
    b = tws.generate_fbm_terrain(N, N, seed=2, octaves=6, base_frequency=4, persistence=0.43,
                                 height_scale=100.0)
    d = np.full((N, N), 0.5, np.float32)
    p = tws.params_physical(**PARAMS)
    sim = tws.Simulation(N, N, p)
    if args.kernel:
        sim.set_launch(args.kernel, args.k or 1)
    kernel, k = sim.get_launch()
    sim.upload(b, d)
    log(f"bench: kernel {kernel}, {k} steps/launch; warmup {args.warmup} steps")
    sim.step(args.warmup)
    sim.sync()

    log(f"bench: timing {args.steps} steps")
    sim.sync()
    t0 = time.perf_counter()
    sim.step(args.steps)
    sim.sync()
    el = time.perf_counter() - t0
    cells = N * N
    value = cells * args.steps / el

    # Per-launch times of the dominant kernel, HIP events on the sim's stream.
    prof_steps = min(args.steps, 200)
    ms = sim.step_profiled(prof_steps)
    full = ms[: prof_steps // k] if k > 1 else ms
    mean_ms = float(np.mean(full))
    bytes_per_launch = BYTES_PER_CELL_UPDATE * cells * k
    achieved = bytes_per_launch / (mean_ms * 1e-3)
    finite = sim.check_finite()
    # The same steps with the boundary's host transfers inside the clock:
    # upload b and d from (pageable) host arrays, step, download d and 4 f.
    sim.sync()
    t0 = time.perf_counter()
    sim.upload(b, d)
    sim.step(args.steps)
    host_out = sim.download(flux=True)
    el_io = time.perf_counter() - t0
    del host_out
    sim.close()
    copy_gbps = tws.stream_copy_gbps(1 << 30, 20)

    traffic = None
    if TRAFFIC_FILE.exists():
        t = json.loads(TRAFFIC_FILE.read_text())
        key = f"{kernel}/{k}/{N}"
        if key in t:
            traffic = t[key]["hbm_bytes_per_launch"]

    rec = {
        "metric": "cell-updates/s at 4096^2 (plain FP32 pipe-model step)",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"plain {N}x{N} FP32 grid, fBm terrain (seed 2, 6 octaves, 0-100 m) "
                        "+ 0.5 m water sheet, closed walls, dt 0.02, g 9.81, damping 0.995",
            "grid": [N, N],
            "kernel": kernel,
            "steps_per_launch": k,
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved / 1e9,
            "peak": HBM_PEAK_BPS / 1e9,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_BPS,
            "traffic": traffic,
            "kernel": "k_tile",
            "mean_launch_ms": mean_ms,
            "algorithmic_bytes_per_launch": bytes_per_launch,
            "stream_copy_GBps_measured": copy_gbps,
        },
        "finite": finite,
        "pcie_inclusive": {
            "cell_updates_per_s": cells * args.steps / el_io,
            "seconds": el_io,
            "compute_only_seconds": el,
            "transfers": "upload b, d from host arrays (flux zeroed on the device), "
                         f"{args.steps} steps, download d and the 4 flux planes",
        },
    }
    if not args.no_cpu_baseline:
        log("bench: CPU baseline (FP32 restatement on host cores)")
        rec["cpu_baseline"] = cpu_baseline(b, d, (p.flow_friction, p.water_acceleration,
                                                  p.cell_area_inv_timescaled),
                                           target_s=args.cpu_seconds)
    del b, d
    extra = [w for w in args.extra.split(",") if w]
    if extra:
        from terrainwatersim_amd import workloads
        if "small" in extra:
            rec["small"] = workloads.small(log=log)
        if "batched" in extra:
            rec["batched"] = workloads.batched(log=log)
        if "served" in extra:
            cap = rec.get("batched", {}).get("batched_cell_updates_per_s", value)
            rec["served"] = workloads.served(cap, n_requests=args.served_requests, log=log)
        if "large" in extra:
            rec["large"] = workloads.large(log=log)
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
