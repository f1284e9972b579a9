#!/usr/bin/env python3
"""Benchmark of the virtual-pipe water step on one MI355X.

Workload (the headline metric's configuration): one 4096 x 4096 FP32 grid,
seeded 6-octave value-noise fBm terrain scaled to 0-100 m plus a uniform
0.5 m water sheet, closed walls, g = 9.81, dt = 0.02 s, cell 1 m, pipe area
1 m^2, damping 0.995.  A bench "step" is one simulation step of the whole
grid; --steps K steps are timed (default 1000), after --warmup W untimed ones.

Printed: one JSON line with cell-updates/s, the roofline of the dominant
kernel and a CPU baseline.

roofline: the step kernel advances k steps per launch with the state held on
chip in between, so the HBM traffic it cannot avoid is one read of b, d and
the 4 fluxes and one write of d and the 4 fluxes per cell and LAUNCH (44 B x
N^2).  `achieved` is that compulsory traffic over the mean launch time (HIP
events on the simulation's stream) and `frac` = achieved / 8 TB/s.  Beside
it: the bytes the PMC counters saw per launch (`traffic`, from the committed
profile named in `traffic_source`), the fraction of VALU issue slots used
(same profile), and the older algorithmic-equivalent figure 44 B x N^2 x k per
launch, which counts the state as crossing HBM every step and so exceeds 1.

cpu_baseline: the FP32 restatement under oracle/ on the host cores, on a
bounded sample of the same grid (kind "port"); `reference_llvmpipe` carries
the reference's own shaders timed on llvmpipe (the committed figure, measured
where the reference checkout is present).  The same leg checks the small
workload's GPU result against the FP64 restatement.
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
LLVMPIPE_FILE = ROOT / "profiles" / "cpu_baseline_llvmpipe.json"
CU_COUNT = 256                 # MI355X compute units
CLOCK_HZ = 2.4e9               # MI355X peak engine clock
KERNEL_STREAM = 8              # TWS_KERNEL_STREAM_128
# one wave64 VALU instruction per cycle per CU (4 SIMDs, 16 lanes each)
VALU_WAVE_INSTS_PER_S = CU_COUNT * CLOCK_HZ


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(b, d, params, target_s=12.0):
    """FP32 restatement on all host cores, on a bounded number of steps of
    the same grid (step count calibrated from two warm steps to about
    target_s seconds), plus the reference shaders' llvmpipe figure."""
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
    out = {
        "value": W * H * steps / el,
        "unit": "cell-updates/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{W}x{H} grid, {steps} steps after 3 warm steps ({el:.1f} s), "
                  "oracle/tws_oracle.c FP32 restatement, OpenMP over rows",
    }
    if LLVMPIPE_FILE.exists():
        ref = json.loads(LLVMPIPE_FILE.read_text())
        big = [r for r in ref["runs"] if r["workload"].startswith("4096")]
        if big:
            r = big[0]
            out["reference_llvmpipe"] = {
                "value": r["cell_updates_per_s"],
                "unit": "cell-updates/s",
                "cores": ref["cores"],
                "kind": "reference",
                "host": ref["cpu"],
                "renderer": ref["renderer"],
                "sample": f"{r['workload']}, {r['steps']} steps in {r['seconds']:.1f} s",
                "where": "unmodified reference shaders, measured in the build container where "
                         "the reference checkout is present (the GPU host has none); "
                         f"{LLVMPIPE_FILE.relative_to(ROOT)}",
                "runs": ref["runs"],
            }
    return out


def small_vs_fp64(small_rec, depth_gpu, params):
    """The small workload (256^2 rain, 1000 steps) against the FP64 and FP32
    restatements on the host: the FP32 error of the GPU result, and whether
    the GPU result equals the FP32 restatement bit for bit."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import tws_oracle
    from terrainwatersim_amd import workloads
    b, d, rain, steps = workloads.small_state()
    F, K, Cc = params
    d32, _, _ = tws_oracle.step(b, d, None, F, K, Cc, n=steps, rain=rain)
    d64, _, _ = tws_oracle.step(b, d, None, F, K, Cc, n=steps, rain=rain, dtype=np.float64)
    err = np.abs(depth_gpu.astype(np.float64) - d64)
    floor = 1e-3 * float(d64.max())
    return {
        "reference": "FP64 restatement (oracle/tws_oracle.c), same initial state and rain",
        "max_abs_m": float(err.max()),
        "max_rel_floor": float((err / np.maximum(d64, floor)).max()),
        "rel_floor_m": floor,
        "max_depth_m": float(d64.max()),
        "gpu_equals_fp32_restatement": bool(np.array_equal(depth_gpu, d32)),
        "volume_fp64_m3": float(np.sum(d64)),
        "volume_gpu_m3": small_rec.get("volume_m3"),
    }


def roofline(kernel, k, N, mean_ms, copy_gbps):
    """Roofline of the step kernel at one launch of k steps (see the module
    docstring)."""
    cells = N * N
    t = mean_ms * 1e-3
    compulsory = BYTES_PER_CELL_UPDATE * cells
    achieved = compulsory / t
    rec = {
        "bound": "hbm",
        "achieved": achieved / 1e9,
        "peak": HBM_PEAK_BPS / 1e9,
        "unit": "GB/s",
        "frac": achieved / HBM_PEAK_BPS,
        "traffic": None,
        "kernel": "k_stream" if kernel == KERNEL_STREAM else "k_tile",
        "steps_per_launch": k,
        "mean_launch_ms": mean_ms,
        "compulsory_bytes_per_launch": compulsory,
        "stream_copy_GBps_measured": copy_gbps,
        "frac_of_copy": achieved / (copy_gbps * 1e9) if copy_gbps else None,
        "algorithmic_equiv": {
            "bytes_per_launch": compulsory * k,
            "GBps": compulsory * k / t / 1e9,
            "frac": compulsory * k / t / HBM_PEAK_BPS,
            "note": "44 B per cell-update, as if the state crossed HBM every step",
        },
    }
    if TRAFFIC_FILE.exists():
        e = json.loads(TRAFFIC_FILE.read_text()).get(f"{kernel}/{k}/{N}")
        if e:
            rec["kernel"] = e.get("kernel_name", rec["kernel"])
            rec["traffic"] = e["hbm_bytes_per_launch"]
            rec["traffic_source"] = e.get("source")
            rec["traffic_over_compulsory"] = e["hbm_bytes_per_launch"] / compulsory
            rec["counted_frac"] = e["hbm_bytes_per_launch"] / t / HBM_PEAK_BPS
            if "valu_insts_per_launch" in e:
                v = e["valu_insts_per_launch"]
                rec["valu"] = {
                    "wave_insts_per_launch": v,
                    "issue_frac": v / (t * VALU_WAVE_INSTS_PER_S),
                    "lane_ops_per_output_cell_update": v * 64 / (cells * k),
                    "source": e.get("source"),
                }
    return rec


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
    copy_gbps = tws.stream_copy_gbps(1 << 30, 10)

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
        "roofline": roofline(kernel, k, N, mean_ms, copy_gbps),
        "finite": finite,
        "pcie_inclusive": {
            "cell_updates_per_s": cells * args.steps / el_io,
            "seconds": el_io,
            "compute_only_seconds": el,
            "transfers": "upload b, d from host arrays (flux zeroed on the device), "
                         f"{args.steps} steps, download d and the 4 flux planes",
        },
    }
    del b
    extra = [w for w in args.extra.split(",") if w]
    small_depth = None
    if extra:
        from terrainwatersim_amd import workloads
        if "small" in extra:
            rec["small"], small_depth = workloads.small(log=log, return_depth=True)
        if "batched" in extra:
            rec["batched"] = workloads.batched(log=log)
        if "served" in extra:
            cap = rec.get("batched", {}).get("batched_cell_updates_per_s", value)
            dev_cap = max(value, rec.get("batched", {}).get("batched_cell_updates_per_s", 0.0))
            rec["served"] = workloads.served(cap, n_requests=args.served_requests, log=log,
                                             device_cap=dev_cap)
        if "large" in extra:
            rec["large"] = workloads.large(log=log)
    if not args.no_cpu_baseline:
        log("bench: CPU baseline (FP32 restatement on host cores)")
        fkc = (p.flow_friction, p.water_acceleration, p.cell_area_inv_timescaled)
        b = tws.generate_fbm_terrain(N, N, seed=2, octaves=6, base_frequency=4,
                                     persistence=0.43, height_scale=100.0)
        rec["cpu_baseline"] = cpu_baseline(b, d, fkc, target_s=args.cpu_seconds)
        del b
        if small_depth is not None:
            log("bench: small workload against the FP64 restatement")
            rec["small"]["fp32_vs_fp64"] = small_vs_fp64(rec["small"], small_depth, fkc)
    del d
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
