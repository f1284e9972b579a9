"""numpy front end of the CPU restatement (oracle/tws_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or the reported CPU
baseline, never as the measured or shipped path.

Functions follow, in the reference (paths from its repository root):
  step()             flowUpdate.comp:14-62 then flowApply.comp:16-52 per step,
                     in Terrain.cpp:253-265 order
  brush()            waterBrush.comp:20-31
  params_reference() Terrain.cpp:175-198
  params_physical()  simulationCommon.glsl:5-12 (the stated meaning of F, K, C)
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "libtws_oracle.so"

BOUNDARY_CLOSED = 0
BOUNDARY_ZERO = 1
BOUNDARY_DISPATCH = 2   # ZERO plus the reference's truncating dispatch shapes

_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        h = C.CDLL(str(LIB))
        vp, i32, f32, f64 = C.c_void_p, C.c_int, C.c_float, C.c_double
        h.oracle_step_f32.argtypes = [i32, i32, vp, vp, vp, f32, f32, f32, i32, f32, vp, i32, vp]
        h.oracle_step_f64.argtypes = [i32, i32, vp, vp, vp, f64, f64, f64, i32, f64, vp, i32, vp]
        h.oracle_step_f32.restype = h.oracle_step_f64.restype = i32
        h.oracle_brush_f32.argtypes = [i32, i32, vp, f32, f32, f32, f32]
        h.oracle_brush_f64.argtypes = [i32, i32, vp, f32, f32, f32, f32]
        h.oracle_params_reference.argtypes = [f32, f32, f32, f32, C.c_uint, vp]
        h.oracle_params_physical.argtypes = [f32, f32, f32, f32, f32, vp]
        h.oracle_sum_f32.argtypes = [vp, C.c_int64]
        h.oracle_sum_f32.restype = f64
        h.oracle_set_threads.argtypes = [i32]
        h.oracle_get_threads.restype = i32
        h.oracle_mip_levels.argtypes = [i32, i32]
        h.oracle_mip_levels.restype = i32
        h.oracle_mips.argtypes = [i32, i32, vp, vp]
        _lib = h
    return _lib


def set_threads(n: int) -> int:
    """Host threads for the restatement; returns the count now in effect."""
    lib().oracle_set_threads(int(n))
    return lib().oracle_get_threads()


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def step(b, d, f=None, F=0.99989975, K=0.1962, Cc=0.02, n=1, boundary=BOUNDARY_CLOSED,
         ghost_alpha=0.0, rain=None, velocity=False, dtype=np.float32):
    """Runs n steps; returns (d, f, v) as new arrays of `dtype` (v is None
    unless velocity).  b, d: (H, W); f: (H, W, 4) or None for zero flux;
    rain: (cx, cy, intensity, radius_sq) applied before every step."""
    dtype = np.dtype(dtype)
    H, W = np.shape(b)
    bb = np.ascontiguousarray(b, dtype=dtype)
    dd = np.array(d, dtype=dtype, copy=True, order="C")
    ff = (np.zeros((H, W, 4), dtype) if f is None
          else np.array(f, dtype=dtype, copy=True, order="C"))
    vv = np.zeros((H, W, 2), dtype) if velocity else None
    rr = None if rain is None else np.asarray(rain, np.float32)
    fn = lib().oracle_step_f32 if dtype == np.float32 else lib().oracle_step_f64
    if dtype == np.float32:
        F, K, Cc, ga = (np.float32(x) for x in (F, K, Cc, ghost_alpha))
    rc = fn(W, H, _p(bb), _p(dd), _p(ff), float(F), float(K), float(Cc), boundary,
            float(ghost_alpha) if dtype != np.float32 else float(ga), _p(rr), int(n), _p(vv))
    if rc != 0:
        raise MemoryError("oracle_step: allocation failed")
    return dd, ff, vv


def brush(d, cx, cy, intensity, radius_sq, dtype=np.float32):
    dd = np.array(d, dtype=dtype, copy=True, order="C")
    H, W = dd.shape
    fn = lib().oracle_brush_f32 if np.dtype(dtype) == np.float32 else lib().oracle_brush_f64
    fn(W, H, _p(dd), cx, cy, intensity, radius_sq)
    return dd


def params_reference(steps_per_second=60.0, damping=0.98, acceleration=10.0,
                     grid_world_size=1024.0, grid_resolution=1024):
    out = np.zeros(3, np.float32)
    lib().oracle_params_reference(steps_per_second, damping, acceleration, grid_world_size,
                                  grid_resolution, _p(out))
    return tuple(float(x) for x in out)


def params_physical(dt=0.02, gravity=9.81, cell_size=1.0, pipe_area=1.0, damping=0.995):
    out = np.zeros(3, np.float32)
    lib().oracle_params_physical(dt, gravity, cell_size, pipe_area, damping, _p(out))
    return tuple(float(x) for x in out)


def volume(d) -> float:
    a = np.ascontiguousarray(d, np.float32)
    return lib().oracle_sum_f32(_p(a), a.size)


def mip_levels(W: int, H: int) -> int:
    """Full mip chain length, Texture.cpp:28-44: floor(log2(max(W, H))) + 1."""
    return int(lib().oracle_mip_levels(int(W), int(H)))


def mips(rgba0):
    """Levels 1.. of the mip chain of an H x W x 4 RGBA32F texture, as the
    list of (h, w, 4) arrays glGenerateMipmap produces on llvmpipe
    (Terrain.cpp:270-276, Texture2D.cpp:64-68)."""
    a = np.ascontiguousarray(rgba0, np.float32)
    H, W, _ = a.shape
    levels = mip_levels(W, H)
    sizes = [(max(1, H >> l), max(1, W >> l)) for l in range(1, levels)]
    out = np.zeros(sum(h * w * 4 for h, w in sizes), np.float32)
    lib().oracle_mips(W, H, _p(a), _p(out))
    res, off = [], 0
    for h, w in sizes:
        res.append(out[off:off + h * w * 4].reshape(h, w, 4))
        off += h * w * 4
    return res
