#!/usr/bin/env python3
"""Writes the golden fixtures of the reference's own GLSL shaders.

Runs terrainwatersim/shader/{flowUpdate,flowApply,waterBrush}.comp, unmodified
and read from the reference checkout at run time, on Mesa llvmpipe through
oracle/_ref/libtws_glref.so (built by `make -C oracle ref`), and stores inputs
and outputs as .npz fixtures plus manifest.json (shader SHA-256s, renderer,
constants as float bit patterns, boundary handling, step counts).

Closed walls are reproduced with the unmodified shaders by surrounding the
domain with a 16-cell frame of dry terrain 1e6 m high: the flux toward the
frame, max(0, f*F + (h - 1e6)*K), is exactly 0, and frame cells (d = 0)
scale their own outflow by 0/S = 0.  The reference-edge mode runs the domain
unpadded, so the shaders read texels outside the image.

The reference-dispatch case runs the unmodified dispatch shapes on a grid
whose sides are not multiples of 16 or 32 (Terrain.cpp:167,258,264 divide
the resolution with truncation), so the fixture records which cells the
reference leaves untouched.

The window case (window_cases in the manifest) pins the headline
configuration directly: the 4096^2 fBm terrain (seed 2) under a 0.5 m sheet,
closed walls, 64 steps; only the centre 256^2 window of the result is
stored, with the SHA-256 of the full terrain it started from.

Also records the llvmpipe timing of the 256^2 rain workload (1000 steps) and
of 20 steps of a 4096^2 grid as profiles/cpu_baseline_llvmpipe.json.

Mip fixtures (mips_*.npz): the TerrainData mip chain that
Terrain::PerformSimulationStep regenerates after each frame with steps
(Terrain.cpp:270-276 -> Texture2D::GenMipMaps, Texture2D.cpp:64-68), i.e.
glGenerateMipmap on a full-chain RGBA32F texture (r = b, g = b = 0.3,
a = d), levels 1.. read back from llvmpipe.

    python tests/golden/make_goldens.py [--shaders DIR] [--skip-timing] [--mips-only]
                                        [--cases NAME,...] [--window]
"""
from __future__ import annotations

import argparse
import ctypes as C
import hashlib
import json
import os
import platform
import subprocess
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
HARNESS = ROOT / "oracle" / "_ref" / "libtws_glref.so"
DRIVER = "/usr/lib/x86_64-linux-gnu/dri/swrast_dri.so"
SHADERS = ["flowUpdate.comp", "flowApply.comp", "waterBrush.comp", "simulationCommon.glsl",
           "helper.glsl", "constantbuffers.glsl"]
F32 = np.float32
FRAME = 16
WALL = 1.0e6

sys.path.insert(0, str(ROOT))


def bits(x: float) -> str:
    return f"0x{np.float32(x).view(np.uint32):08x}"


class GLRef:
    def __init__(self, shader_dir: str):
        if not HARNESS.exists():
            subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "ref"], check=True)
        os.environ.setdefault("GALLIUM_DRIVER", "llvmpipe")
        self.lib = C.CDLL(str(HARNESS))
        self.lib.glref_last_error.restype = C.c_char_p
        self.lib.glref_renderer.restype = C.c_char_p
        vp = C.c_void_p
        self.lib.glref_run.argtypes = [C.c_int, C.c_int, vp, vp, C.c_float, C.c_float, C.c_float,
                                       vp, C.c_int, vp, vp, vp, C.POINTER(C.c_double)]
        if self.lib.glref_init(shader_dir.encode(), DRIVER.encode()) != 0:
            raise RuntimeError(self.lib.glref_last_error().decode())
        self.renderer = self.lib.glref_renderer().decode()

    def run(self, b, d, f, FKC, n, rain=None):
        H, W = b.shape
        t = np.zeros((H, W, 4), F32)
        t[..., 0], t[..., 1], t[..., 2], t[..., 3] = b, 0.3, 0.3, d
        ff = np.ascontiguousarray(f if f is not None else np.zeros((H, W, 4)), F32)
        ot = np.zeros_like(t)
        of = np.zeros((H, W, 4), F32)
        om = np.zeros((H, W, 2), np.uint16)
        r = None if rain is None else np.asarray(rain, F32)
        sec = C.c_double()
        p = lambda a: None if a is None else a.ctypes.data_as(C.c_void_p)
        rc = self.lib.glref_run(W, H, p(t), p(ff), FKC[0], FKC[1], FKC[2], p(r), int(n), p(ot),
                                p(of), p(om), C.byref(sec))
        if rc != 0:
            raise RuntimeError(self.lib.glref_last_error().decode())
        return ot[..., 3].copy(), of, om, sec.value

    def mips(self, b, d, gb=0.3):
        """Levels 1.. of glGenerateMipmap over the RGBA32F texture (b, gb, gb, d)."""
        H, W = b.shape
        t = np.zeros((H, W, 4), F32)
        t[..., 0], t[..., 1], t[..., 2], t[..., 3] = b, gb, gb, d
        levels = 0
        w, h = W, H
        while w > 0 or h > 0:
            w, h, levels = w // 2, h // 2, levels + 1
        sizes = [(max(1, H >> l), max(1, W >> l)) for l in range(1, levels)]
        out = np.zeros(sum(hh * ww * 4 for hh, ww in sizes), F32)
        sec = C.c_double()
        fn = self.lib.glref_genmips
        fn.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.POINTER(C.c_double)]
        if fn(W, H, levels, t.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p),
              C.byref(sec)) != 0:
            raise RuntimeError(self.lib.glref_last_error().decode())
        res, off = [], 0
        for hh, ww in sizes:
            res.append(out[off:off + hh * ww * 4].reshape(hh, ww, 4).copy())
            off += hh * ww * 4
        return res


def framed(b, d, f):
    H, W = b.shape
    bp = np.full((H + 2 * FRAME, W + 2 * FRAME), WALL, F32)
    dp = np.zeros_like(bp)
    fp = np.zeros(bp.shape + (4,), F32)
    bp[FRAME:-FRAME, FRAME:-FRAME] = b
    dp[FRAME:-FRAME, FRAME:-FRAME] = d
    if f is not None:
        fp[FRAME:-FRAME, FRAME:-FRAME] = f
    return bp, dp, fp


def interior(a):
    return a[FRAME:-FRAME, FRAME:-FRAME]


def run_case(gl, b, d, f, FKC, checkpoints, boundary, rain=None):
    out = {}
    for n in checkpoints:
        if boundary == "closed":
            bp, dp, fp = framed(b, d, f)
            rr = None if rain is None else (rain[0] + FRAME, rain[1] + FRAME, rain[2], rain[3])
            gd, gf, gm, _ = gl.run(bp, dp, fp, FKC, n, rr)
            frame = np.ones(gd.shape, bool)
            frame[FRAME:-FRAME, FRAME:-FRAME] = False
            assert not gd[frame].any() and not gf[frame].any(), "frame took water"
            gd, gf, gm = interior(gd), interior(gf), interior(gm)
        else:
            gd, gf, gm, _ = gl.run(b, d, f, FKC, n, rain)
        out[f"d_{n}"], out[f"f_{n}"], out[f"v16_{n}"] = gd, gf, gm
    return out


def write_mips(gl, tws):
    """Mip-chain fixtures: power-of-two shapes (llvmpipe's blit halves every
    level exactly) and non-power-of-two ones (its sample positions carry the
    rasteriser's rounding)."""
    rng = np.random.default_rng(7)
    cases = {
        "mips_fbm128": (tws.generate_fbm_terrain(128, 128, seed=7),
                        (rng.random((128, 128)) * 3.0 * (rng.random((128, 128)) < 0.6)).astype(F32)),
        "mips_rect64x16": ((rng.standard_normal((16, 64)) * 50).astype(F32),
                           (rng.random((16, 64)) * 2).astype(F32)),
        "mips_column1x8": ((rng.standard_normal((8, 1)) * 50).astype(F32),
                           (rng.random((8, 1)) * 2).astype(F32)),
        "mips_npot100x60": (tws.generate_fbm_terrain(100, 60, seed=8),
                            (rng.random((60, 100)) * 3.0).astype(F32)),
    }
    entries = {}
    for name, (b, d) in cases.items():
        levels = gl.mips(b, d)
        arrays = {"b": b, "d": d}
        arrays.update({f"level{l}": a for l, a in enumerate(levels, 1)})
        np.savez_compressed(HERE / f"{name}.npz", **arrays)
        entries[name] = {"shape": list(b.shape), "gb_fill": 0.3, "levels": len(levels) + 1,
                         "arrays": {k: hashlib.sha256(np.ascontiguousarray(v).tobytes()).hexdigest()
                                    for k, v in arrays.items()}}
        print(f"{name}: {len(levels)} levels", flush=True)
    return entries


WINDOW = 256


def write_window(gl, tws, FKC, N=4096, steps=64):
    """The headline configuration (4096^2 fBm seed 2 + 0.5 m sheet, closed
    walls) after `steps` steps of the shaders; the centre WINDOW^2 of d, f
    and the FlowMap bits is stored."""
    b = tws.generate_fbm_terrain(N, N, seed=2, octaves=6, base_frequency=4, persistence=0.43,
                                 height_scale=100.0)
    d = np.full((N, N), 0.5, F32)
    t0 = time.time()
    out = run_case(gl, b, d, None, FKC, (steps,), "closed")
    y0 = x0 = (N - WINDOW) // 2
    win = lambda a: np.ascontiguousarray(a[y0:y0 + WINDOW, x0:x0 + WINDOW])
    arrays = {"b_window": win(b), f"d_{steps}": win(out[f"d_{steps}"]),
              f"f_{steps}": win(out[f"f_{steps}"]), f"v16_{steps}": win(out[f"v16_{steps}"])}
    name = f"config2_{N}_window{steps}"
    np.savez_compressed(HERE / f"{name}.npz", **arrays)
    print(f"{name}: {time.time() - t0:.1f}s", flush=True)
    return {name: {
        "shape": [N, N], "boundary": "closed", "steps": [steps],
        "initial": "terrain generate_fbm_terrain(4096, 4096, seed=2, octaves=6, base_frequency=4, "
                   "persistence=0.43, height_scale=100); depth 0.5 m everywhere; zero flux",
        "terrain_sha256": hashlib.sha256(b.tobytes()).hexdigest(),
        "window": [y0, x0, WINDOW, WINDOW],
        "F": bits(FKC[0]), "K": bits(FKC[1]), "C": bits(FKC[2]),
        "F_value": float(FKC[0]), "K_value": float(FKC[1]), "C_value": float(FKC[2]),
        "arrays": {k: hashlib.sha256(np.ascontiguousarray(v).tobytes()).hexdigest()
                   for k, v in arrays.items()},
    }}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--shaders", default=os.environ.get(
        "TWS_REFERENCE_SHADERS", "/root/reference/terrainwatersim/shader"))
    ap.add_argument("--skip-timing", action="store_true")
    ap.add_argument("--mips-only", action="store_true",
                    help="only (re)write the mip fixtures and their manifest entries")
    ap.add_argument("--cases", default="",
                    help="comma list: only (re)write these step cases (others kept as they are)")
    ap.add_argument("--window", action="store_true",
                    help="only (re)write the 4096^2 centre-window case")
    args = ap.parse_args()

    import terrainwatersim_amd as tws   # host-only functions: params, terrain generator

    gl = GLRef(args.shaders)
    print("renderer:", gl.renderer, flush=True)
    phys = tws.params_physical(0.02, 9.81, 1.0, 1.0, 0.995)
    FKC = (phys.flow_friction, phys.water_acceleration, phys.cell_area_inv_timescaled)
    refp = tws.params_reference(60.0, 0.98, 10.0, 1024.0, 1024)
    FKC_ref = (refp.flow_friction, refp.water_acceleration, refp.cell_area_inv_timescaled)
    manifest = {
        "generator": "tests/golden/make_goldens.py",
        "renderer": gl.renderer,
        "shaders": {s: hashlib.sha256((Path(args.shaders) / s).read_bytes()).hexdigest()
                    for s in SHADERS},
        "closed_walls": f"{FRAME}-cell frame of dry terrain {WALL:g} m high around the domain",
        "cases": {},
    }

    if args.mips_only:
        old = json.loads((HERE / "manifest.json").read_text())
        old["mip_cases"] = write_mips(gl, tws)
        (HERE / "manifest.json").write_text(json.dumps(old, indent=1) + "\n")
        return
    if args.window:
        old = json.loads((HERE / "manifest.json").read_text())
        old.setdefault("window_cases", {}).update(write_window(gl, tws, FKC))
        (HERE / "manifest.json").write_text(json.dumps(old, indent=1) + "\n")
        return

    def rnd(seed, H, W, wet=0.7):
        rng = np.random.default_rng(seed)
        b = (rng.random((H, W)) * 5.0).astype(F32)
        d = (rng.random((H, W)) * 2.0 * (rng.random((H, W)) < wet)).astype(F32)
        f = (rng.random((H, W, 4)) * 0.5).astype(F32)
        return b, d, f

    cases = []
    # 1. the 256^2 rain workload: fBm terrain seed 1, dry start, rain disc
    #    r = 16 (R^2 = 256) adding 1e-3 m/step at the centre, closed walls
    b = tws.generate_fbm_terrain(256, 256, seed=1)
    cases.append(("config1_256_rain", b, np.zeros((256, 256), F32), None, FKC, (1, 1000),
                  "closed", (128.0, 128.0, 1e-3, 256.0)))
    # 2. random state, rain, closed walls
    b, d, f = rnd(101, 64, 64)
    cases.append(("random64_closed_rain", b, d, f, FKC, (1, 37), "closed", (40.0, 30.0, 1e-2, 256.0)))
    # 3. random state, the reference's own edge (texels outside read as 0)
    b, d, f = rnd(102, 64, 96)
    cases.append(("random64x96_zero_edge", b, d, f, FKC, (1, 20), "zero", None))
    # 4. the reference's default constants (60 steps/s, damping 0.98, accel 10)
    b, d, f = rnd(103, 64, 64, wet=0.4)
    cases.append(("random64_reference_defaults", b, d, f, FKC_ref, (50,), "closed", None))
    # 5. denormal depths and fluxes
    rng = np.random.default_rng(104)
    b = (rng.random((32, 32)) * 1e-3).astype(F32)
    d = (rng.random((32, 32)) * 1e-38).astype(F32)
    f = (rng.random((32, 32, 4)) * 1e-39).astype(F32)
    cases.append(("denormal32", b, d, f, FKC, (1, 5), "closed", None))
    # 6. the reference's own edge at a size its dispatch shapes truncate:
    #    100 x 88 runs 6 x 5 work groups of 16^2 (columns >= 96 and rows >= 80
    #    never change) and a 3 x 2 brush grid of 32^2 (rain only below row 64)
    b, d, f = rnd(106, 88, 100)
    cases.append(("random100x88_dispatch", b, d, f, FKC, (1, 25), "dispatch",
                  (50.0, 70.0, 1e-2, 400.0)))

    only = {c for c in args.cases.split(",") if c}
    if only:
        old = json.loads((HERE / "manifest.json").read_text())
        manifest["cases"] = old["cases"]
        manifest["mip_cases"] = old.get("mip_cases", {})
        manifest["window_cases"] = old.get("window_cases", {})
        cases = [c for c in cases if c[0] in only]
    for name, b, d, f, fkc, cps, boundary, rain in cases:
        t0 = time.time()
        out = run_case(gl, b, d, f, fkc, cps, boundary, rain)
        arrays = {"b": b, "d0": d}
        if f is not None:
            arrays["f0"] = f
        arrays.update(out)
        np.savez_compressed(HERE / f"{name}.npz", **arrays)
        manifest["cases"][name] = {
            "shape": list(b.shape), "boundary": boundary, "steps": list(cps),
            "F": bits(fkc[0]), "K": bits(fkc[1]), "C": bits(fkc[2]),
            "F_value": float(fkc[0]), "K_value": float(fkc[1]), "C_value": float(fkc[2]),
            "rain": None if rain is None else [float(x) for x in rain],
            "arrays": {k: hashlib.sha256(np.ascontiguousarray(v).tobytes()).hexdigest()
                       for k, v in arrays.items()},
        }
        print(f"{name}: {time.time() - t0:.1f}s", flush=True)
    if only:
        (HERE / "manifest.json").write_text(json.dumps(manifest, indent=1) + "\n")
        return
    manifest["mip_cases"] = write_mips(gl, tws)
    manifest["window_cases"] = write_window(gl, tws, FKC)
    (HERE / "manifest.json").write_text(json.dumps(manifest, indent=1) + "\n")

    if args.skip_timing:
        return
    cpu = platform.processor() or "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    timing = {"renderer": gl.renderer, "cpu": cpu, "cores": os.cpu_count(),
              "LP_NUM_THREADS": os.environ.get("LP_NUM_THREADS", "default (all cores)"),
              "what": "unmodified reference shaders on llvmpipe; glFinish + steady_clock "
                      "around the step loop (brush + flowUpdate + flowApply per step)",
              "runs": []}
    b = tws.generate_fbm_terrain(256, 256, seed=1)
    bp, dp, fp = framed(b, np.zeros((256, 256), F32), None)
    _, _, _, sec = gl.run(bp, dp, fp, FKC, 1000, (128.0 + FRAME, 128.0 + FRAME, 1e-3, 256.0))
    timing["runs"].append({"workload": "256x256 fBm, rain, closed walls (framed 288x288 texture)",
                           "steps": 1000, "seconds": sec,
                           "cell_updates_per_s": 256 * 256 * 1000 / sec})
    N = 4096
    b = tws.generate_fbm_terrain(N, N, seed=2)
    d = np.full((N, N), 0.5, F32)
    gl.run(b, d, None, FKC, 1)  # warm
    _, _, _, sec = gl.run(b, d, None, FKC, 20)
    timing["runs"].append({"workload": "4096x4096 fBm + 0.5 m sheet (reference edge, unpadded)",
                           "steps": 20, "seconds": sec, "cell_updates_per_s": N * N * 20 / sec,
                           "note": "1000-step time extrapolated: %.1f s" % (sec * 50)})
    (ROOT / "profiles" / "cpu_baseline_llvmpipe.json").write_text(json.dumps(timing, indent=1) + "\n")
    print(json.dumps(timing, indent=1))


if __name__ == "__main__":
    main()
