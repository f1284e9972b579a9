"""terrainwatersim_amd -- MI355X-native virtual-pipe water simulation.

The product is libtws.so (HIP kernels for gfx950 + C ABI, include/tws.h) and
the C++ host headers over it.  This Python module is a thin ctypes binding
used by tests and the benchmark; it moves numpy arrays across the ABI and
does no simulation work itself.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import (BATCH_MAX_CLASSES, BOUNDARY_CLOSED, BOUNDARY_REFERENCE_DISPATCH, BOUNDARY_ZERO, KERNEL_AUTO, KERNEL_TILE_128x64,
                   KERNEL_TILE_128x128, KERNEL_TILE_256x64, KERNEL_TILE_64x64,
                   KERNEL_TILE_64x64_16W, KERNEL_TILE_64x48_16W, KERNEL_STREAM_128, KERNEL_TWOPASS, LIB_PATH,
                   PROTOTYPES, REQ_CANCELLED, REQ_DONE, REQ_FAILED, REQ_QUEUED, REQ_RUNNING,
                   TILE_KERNELS, TWS_ERR_CAPACITY, BatchStats, DevicePlanes, Params, Request,
                   RequestStatus, ServerConfig, ServerStats, TuneEntry, TwsError, check, lib)

__all__ = [
    "Simulation", "Batch", "Server", "Params", "TwsError", "grid_footprint", "params_physical", "params_reference",
    "generate_fbm_terrain", "generate_reference_scene", "device_count", "lane_shift_probe", "lib", "LIB_PATH", "PROTOTYPES",
    "BOUNDARY_CLOSED", "BOUNDARY_ZERO", "BOUNDARY_REFERENCE_DISPATCH", "KERNEL_AUTO", "KERNEL_TWOPASS", "KERNEL_TILE_128x128",
    "KERNEL_TILE_128x64", "KERNEL_TILE_256x64", "KERNEL_TILE_64x64", "KERNEL_TILE_64x64_16W",
    "KERNEL_TILE_64x48_16W", "KERNEL_STREAM_128",
    "TILE_KERNELS", "BATCH_MAX_CLASSES",
]


def _ptr(a: np.ndarray | None):
    if a is None:
        return None
    return a.ctypes.data_as(C.c_void_p)


def _f32(a, shape) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.float32)
    if a.shape != shape:
        raise ValueError(f"expected shape {shape}, got {a.shape}")
    return a


def params_physical(dt=0.02, gravity=9.81, cell_size=1.0, pipe_area=1.0, damping=0.995,
                    boundary=BOUNDARY_CLOSED) -> Params:
    p = Params()
    check(lib().tws_params_physical(dt, gravity, cell_size, pipe_area, damping, boundary,
                                    C.byref(p)))
    return p


def params_reference(steps_per_second=60.0, damping=0.98, acceleration=10.0,
                     grid_world_size=1024.0, grid_resolution=1024,
                     boundary=BOUNDARY_CLOSED) -> Params:
    p = Params()
    check(lib().tws_params_reference(steps_per_second, damping, acceleration, grid_world_size,
                                     grid_resolution, boundary, C.byref(p)))
    return p


def params_raw(F: float, K: float, Cc: float, boundary=BOUNDARY_CLOSED,
               ghost_alpha: float = 0.0) -> Params:
    p = Params()
    p.flow_friction, p.water_acceleration, p.cell_area_inv_timescaled = F, K, Cc
    p.boundary, p.ghost_alpha = boundary, ghost_alpha
    return p


def generate_reference_scene(resolution: int = 1024, height_scale: float = 300.0,
                             seed: int = 231656522, scene_index: int = 0):
    """The reference's own default scene (terrain b, lake depth d), bit for bit:
    tws_generate_reference_scene.  Returns two (resolution, resolution) arrays."""
    b = np.empty((resolution, resolution), np.float32)
    d = np.empty((resolution, resolution), np.float32)
    check(lib().tws_generate_reference_scene(int(resolution), float(height_scale), int(seed),
                                             int(scene_index), _ptr(b), _ptr(d)))
    return b, d


def generate_fbm_terrain(width: int, height: int, seed: int, octaves: int = 6,
                         base_frequency: int = 4, persistence: float = 0.43,
                         height_scale: float = 100.0) -> np.ndarray:
    out = np.empty((height, width), np.float32)
    check(lib().tws_generate_fbm_terrain(width, height, seed, octaves, base_frequency,
                                         persistence, height_scale, _ptr(out)))
    return out


def device_count() -> int:
    return lib().tws_device_count()


def lane_shift_probe() -> np.ndarray:
    out = (C.c_int32 * 128)()
    check(lib().tws_debug_lane_shift_probe(out))
    return np.frombuffer(out, dtype=np.int32).copy()


def float_to_half_rtz(a) -> np.ndarray:
    """float32 -> float16 bits, rounded toward zero (the FlowMap rounding)."""
    x = np.ascontiguousarray(a, np.float32)
    out = np.empty(x.shape, np.uint16)
    check(lib().tws_float_to_half_rtz(_ptr(x), _ptr(out), x.size))
    return out


def mip_chain_size(width: int, height: int):
    """(levels, floats of levels 1..) of the terrain mip chain (Texture.cpp:28-44)."""
    n, f = C.c_int32(), C.c_size_t()
    check(lib().tws_mip_chain_size(int(width), int(height), C.byref(n), C.byref(f)))
    return n.value, f.value


def stream_copy_gbps(nbytes: int = 1 << 30, iterations: int = 20, shape: int = -1) -> float:
    """Measured device-to-device copy bandwidth (read + write bytes / s, GB/s)
    in one launch shape, or the best of all shapes (shape = -1)."""
    g = C.c_double()
    check(lib().tws_stream_copy_bench(int(nbytes), int(iterations), int(shape), C.byref(g)))
    return g.value


def stream_copy_shapes() -> int:
    """Number of launch shapes of the copy kernel."""
    return int(lib().tws_stream_copy_shapes())


class Simulation:
    """One grid on the device (mirrors the simulation half of the reference's
    Terrain class, terrainwatersim/source/scene/Terrain.h:21-39,75-83)."""

    def __init__(self, width: int, height: int, params: Params | None = None, stream=None):
        self.width, self.height = int(width), int(height)
        self._h = C.c_void_p()
        p = params if params is not None else params_physical()
        check(lib().tws_create(self.width, self.height, C.byref(p), stream, C.byref(self._h)))

    # -- lifetime
    def close(self):
        if self._h:
            lib().tws_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- state
    def upload(self, terrain=None, depth=None, flux=None):
        shp = (self.height, self.width)
        t = _f32(terrain, shp) if terrain is not None else None
        d = _f32(depth, shp) if depth is not None else None
        f = _f32(flux, shp + (4,)) if flux is not None else None
        check(lib().tws_upload(self._h, _ptr(t), _ptr(d), _ptr(f)))

    def download(self, depth=True, flux=True, velocity=False):
        shp = (self.height, self.width)
        d = np.empty(shp, np.float32) if depth else None
        f = np.empty(shp + (4,), np.float32) if flux else None
        v = np.empty(shp + (2,), np.float32) if velocity else None
        check(lib().tws_download(self._h, _ptr(d), _ptr(f), _ptr(v)))
        return {"depth": d, "flux": f, "velocity": v}

    def upload_reference_layout(self, terrain_rgba=None, flow_rgba=None):
        shp = (self.height, self.width, 4)
        t = _f32(terrain_rgba, shp) if terrain_rgba is not None else None
        f = _f32(flow_rgba, shp) if flow_rgba is not None else None
        check(lib().tws_upload_reference_layout(self._h, _ptr(t), _ptr(f)))

    def download_reference_layout(self, gb_fill=0.3):
        shp = (self.height, self.width)
        t = np.empty(shp + (4,), np.float32)
        f = np.empty(shp + (4,), np.float32)
        m = np.empty(shp + (2,), np.uint16)
        check(lib().tws_download_reference_layout(self._h, _ptr(t), gb_fill, _ptr(f), _ptr(m)))
        return {"terrain_rgba": t, "flow_rgba": f, "flowmap_rg16f": m}

    # -- parameters and sources
    def set_params(self, p: Params):
        check(lib().tws_set_params(self._h, C.byref(p)))

    def get_params(self) -> Params:
        p = Params()
        check(lib().tws_get_params(self._h, C.byref(p)))
        return p

    def apply_brush(self, cx, cy, intensity, radius_sq):
        check(lib().tws_apply_brush(self._h, cx, cy, intensity, radius_sq))

    def set_rain(self, cx, cy, intensity, radius_sq):
        check(lib().tws_set_rain(self._h, cx, cy, intensity, radius_sq))

    # -- stepping
    def step(self, n: int = 1):
        check(lib().tws_step(self._h, int(n)))

    def step_profiled(self, n: int, capacity: int | None = None) -> np.ndarray:
        """Steps n and returns the duration (ms) of each kernel launch."""
        cap = capacity or max(1, n)
        buf = (C.c_float * cap)()
        nl = C.c_int32()
        check(lib().tws_step_profiled(self._h, int(n), buf, cap, C.byref(nl)))
        return np.frombuffer(buf, dtype=np.float32)[:min(nl.value, cap)].copy()

    def perform_simulation_step(self, frame_seconds: float, step_seconds: float) -> int:
        n = C.c_int32()
        check(lib().tws_perform_simulation_step(self._h, frame_seconds, step_seconds, C.byref(n)))
        return n.value

    # -- terrain mip chain (Terrain.cpp:270-276)
    def mips(self, gb_fill=0.3):
        """Levels 1.. of the TerrainData mip chain as (h, w, 4) float32 arrays."""
        levels, floats = mip_chain_size(self.width, self.height)
        out = np.empty(max(floats, 1), np.float32)
        check(lib().tws_download_mips(self._h, gb_fill, _ptr(out), floats))
        res, off = [], 0
        for lv in range(1, levels):
            h, w = max(1, self.height >> lv), max(1, self.width >> lv)
            res.append(out[off:off + h * w * 4].reshape(h, w, 4))
            off += h * w * 4
        return res

    def generate_mips(self, device_ptr: int, capacity_floats: int, gb_fill=0.3):
        """Writes levels 1.. into device memory at device_ptr (async)."""
        check(lib().tws_generate_mips(self._h, gb_fill, C.c_void_p(device_ptr), capacity_floats))

    def set_frame_mips(self, device_ptr: int | None, capacity_floats: int = 0, gb_fill=0.3):
        """Device buffer perform_simulation_step keeps current (None: off)."""
        check(lib().tws_set_frame_mips(self._h, C.c_void_p(device_ptr) if device_ptr else None,
                                       capacity_floats, gb_fill))

    def set_launch(self, kernel: int, steps_per_launch: int):
        check(lib().tws_set_launch(self._h, kernel, steps_per_launch))

    def get_launch(self):
        k, s = C.c_int32(), C.c_int32()
        check(lib().tws_get_launch(self._h, C.byref(k), C.byref(s)))
        return k.value, s.value

    def init_terrain_fbm(self, seed: int, octaves: int = 6, base_frequency: int = 4,
                         persistence: float = 0.43, height_scale: float = 100.0):
        check(lib().tws_init_terrain_fbm(self._h, int(seed), octaves, base_frequency, persistence,
                                         height_scale))

    def init_water(self, depth: float = 0.0, rect=(0, 0, 0, 0), rect_depth: float = 0.0):
        x0, y0, x1, y1 = rect
        check(lib().tws_init_water(self._h, depth, x0, y0, x1, y1, rect_depth))

    def init_reference_scene(self, height_scale: float = 300.0, seed: int = 231656522,
                             scene_index: int = 0):
        """The reference's default scene computed on the device (the bits of
        generate_reference_scene); square grids of at least 16 x 16."""
        check(lib().tws_init_reference_scene(self._h, float(height_scale), int(seed),
                                             int(scene_index)))

    def set_tile_order(self, group_columns: int):
        check(lib().tws_set_tile_order(self._h, int(group_columns)))

    def sync(self):
        check(lib().tws_sync(self._h))

    def volume(self) -> float:
        v = C.c_double()
        check(lib().tws_volume(self._h, C.byref(v)))
        return v.value

    def check_finite(self) -> bool:
        code = lib().tws_check_finite(self._h)
        if code == -4:
            return False
        check(code)
        return True

    def device_planes(self) -> DevicePlanes:
        p = DevicePlanes()
        check(lib().tws_get_device_planes(self._h, C.byref(p)))
        return p


def tune_record(width: int, height: int, kernel: int, steps_per_launch: int):
    """Makes TWS_KERNEL_AUTO pick (kernel, k) for grids of this size class."""
    check(lib().tws_tune_record(int(width), int(height), int(kernel), int(steps_per_launch)))


def autotune(shapes, steps: int = 100):
    """Times every (variant, k) candidate on scratch grids of `shapes`
    [(height, width), ...]; returns (best, all trials) as dicts.  Records
    nothing (tune_record does)."""
    hs = np.ascontiguousarray([h for h, w in shapes], np.int32)
    ws = np.ascontiguousarray([w for h, w in shapes], np.int32)
    best = TuneEntry()
    cap = 64
    res = (TuneEntry * cap)()
    cnt = C.c_int32()
    check(lib().tws_autotune(_ptr(ws), _ptr(hs), len(shapes), int(steps), C.byref(best), res, cap,
                             C.byref(cnt)))
    conv = lambda e: {"kernel": e.kernel, "k": e.steps_per_launch,
                      "cell_updates_per_s": e.cell_updates_per_s}
    return conv(best), [conv(res[i]) for i in range(cnt.value)]


def tune_forget():
    check(lib().tws_tune_forget())


def grid_footprint(width: int, height: int) -> int:
    """Device bytes of one grid's state (its admission footprint in a batch)."""
    return lib().tws_grid_footprint(int(width), int(height))


class Batch:
    """Many independent grids advanced by one launch sequence (tws_batch_*)."""

    def __init__(self, max_grids: int, pool_bytes: int, stream=None):
        self._h = C.c_void_p()
        self.shapes = {}
        check(lib().tws_batch_create(int(max_grids), int(pool_bytes), stream, C.byref(self._h)))

    def close(self):
        if self._h:
            lib().tws_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, width: int, height: int, params: Params | None = None) -> int:
        gid = C.c_int32()
        p = params if params is not None else params_physical()
        check(lib().tws_batch_add(self._h, int(width), int(height), C.byref(p), C.byref(gid)))
        self.shapes[gid.value] = (int(height), int(width))
        return gid.value

    def try_add(self, width, height, params=None):
        """Like add(), but returns None when the grid does not fit."""
        try:
            return self.add(width, height, params)
        except TwsError as e:
            if e.code == TWS_ERR_CAPACITY:
                return None
            raise

    def remove(self, gid: int):
        check(lib().tws_batch_remove(self._h, int(gid)))
        self.shapes.pop(gid, None)

    def upload(self, gid, terrain=None, depth=None, flux=None):
        shp = self.shapes[gid]
        t = _f32(terrain, shp) if terrain is not None else None
        d = _f32(depth, shp) if depth is not None else None
        f = _f32(flux, shp + (4,)) if flux is not None else None
        check(lib().tws_batch_upload(self._h, int(gid), _ptr(t), _ptr(d), _ptr(f)))

    def download(self, gid, depth=True, flux=True, velocity=False):
        shp = self.shapes[gid]
        d = np.empty(shp, np.float32) if depth else None
        f = np.empty(shp + (4,), np.float32) if flux else None
        v = np.empty(shp + (2,), np.float32) if velocity else None
        check(lib().tws_batch_download(self._h, int(gid), _ptr(d), _ptr(f), _ptr(v)))
        return {"depth": d, "flux": f, "velocity": v}

    def set_params(self, gid, p: Params):
        check(lib().tws_batch_set_params(self._h, int(gid), C.byref(p)))

    def init_terrain_fbm(self, gid, seed: int, octaves: int = 6, base_frequency: int = 4,
                         persistence: float = 0.43, height_scale: float = 100.0):
        check(lib().tws_batch_init_terrain_fbm(self._h, int(gid), int(seed), octaves,
                                               base_frequency, persistence, height_scale))

    def init_water(self, gid, depth: float = 0.0, rect=(0, 0, 0, 0), rect_depth: float = 0.0):
        x0, y0, x1, y1 = rect
        check(lib().tws_batch_init_water(self._h, int(gid), depth, x0, y0, x1, y1, rect_depth))

    def init_reference_scene(self, gid, height_scale: float = 300.0, seed: int = 231656522,
                             scene_index: int = 0):
        check(lib().tws_batch_init_reference_scene(self._h, int(gid), float(height_scale),
                                                   int(seed), int(scene_index)))

    def set_rain(self, gid, cx, cy, intensity, radius_sq):
        check(lib().tws_batch_set_rain(self._h, int(gid), cx, cy, intensity, radius_sq))

    def volume(self, gid) -> float:
        v = C.c_double()
        check(lib().tws_batch_volume(self._h, int(gid), C.byref(v)))
        return v.value

    def check_finite(self, gid) -> bool:
        code = lib().tws_batch_check_finite(self._h, int(gid))
        if code == -4:
            return False
        check(code)
        return True

    def set_steps(self, gid, steps: int):
        check(lib().tws_batch_set_steps(self._h, int(gid), int(steps)))

    def run(self):
        check(lib().tws_batch_run(self._h))

    def step(self, n: int):
        check(lib().tws_batch_step(self._h, int(n)))

    def set_launch(self, kernel: int, steps_per_launch: int):
        check(lib().tws_batch_set_launch(self._h, kernel, steps_per_launch))

    def get_launch(self):
        k, s = C.c_int32(), C.c_int32()
        check(lib().tws_batch_get_launch(self._h, C.byref(k), C.byref(s)))
        return k.value, s.value

    def set_launch_classes(self, classes):
        """Size classes [(max_cells, kernel, steps_per_launch), ...], ascending
        by max_cells; the last class takes every larger grid."""
        n = len(classes)
        mc = np.array([int(c[0]) for c in classes], np.int64)
        kern = np.array([int(c[1]) for c in classes], np.int32)
        ks = np.array([int(c[2]) for c in classes], np.int32)
        check(lib().tws_batch_set_launch_classes(self._h, n, mc.ctypes.data, kern.ctypes.data,
                                                 ks.ctypes.data))

    def get_launch_classes(self):
        n = C.c_int32()
        mc = np.zeros(BATCH_MAX_CLASSES, np.int64)
        kern = np.zeros(BATCH_MAX_CLASSES, np.int32)
        ks = np.zeros(BATCH_MAX_CLASSES, np.int32)
        check(lib().tws_batch_get_launch_classes(self._h, BATCH_MAX_CLASSES, mc.ctypes.data,
                                                 kern.ctypes.data, ks.ctypes.data, C.byref(n)))
        return [(int(mc[c]), int(kern[c]), int(ks[c])) for c in range(n.value)]

    def use_graphs(self, enable: bool):
        check(lib().tws_batch_use_graphs(self._h, 1 if enable else 0))

    def stats(self) -> dict:
        st = BatchStats()
        check(lib().tws_batch_stats_get(self._h, C.byref(st)))
        return {name: getattr(st, name) for name, _ in st._fields_}

    def sync(self):
        check(lib().tws_batch_sync(self._h))


class Server:
    """Continuous-batching request loop (tws_server_*).  Inputs are copied at
    submit; output arrays are kept alive here until the request finishes."""

    def __init__(self, pool_bytes: int, max_resident: int = 64, chunk_steps: int = 48,
                 kernel: int = 0, steps_per_launch: int = 0, max_queue: int = 0):
        cfg = ServerConfig(int(pool_bytes), int(max_resident), int(chunk_steps), int(kernel),
                           int(steps_per_launch), int(max_queue))
        self._h = C.c_void_p()
        self._keep = {}
        check(lib().tws_server_create(C.byref(cfg), C.byref(self._h)))

    def close(self):
        if self._h:
            lib().tws_server_destroy(self._h)
            self._h = C.c_void_p()
            self._keep.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, terrain, steps, params=None, depth=None, flux=None, rain=None,
               want_flux=True, want_velocity=False):
        """Returns (ticket, outputs) where outputs holds the arrays the server
        fills when the request completes."""
        H, W = np.shape(terrain)
        t = _f32(terrain, (H, W))
        d = _f32(depth, (H, W)) if depth is not None else None
        f = _f32(flux, (H, W, 4)) if flux is not None else None
        out = {"depth": np.zeros((H, W), np.float32),
               "flux": np.zeros((H, W, 4), np.float32) if want_flux else None,
               "velocity": np.zeros((H, W, 2), np.float32) if want_velocity else None}
        r = Request()
        r.width, r.height, r.steps = W, H, int(steps)
        r.params = params if params is not None else params_physical()
        r.terrain, r.depth, r.flux4 = _ptr(t), _ptr(d), _ptr(f)
        if rain is not None:
            r.rain_on = 1
            for i in range(4):
                r.rain[i] = rain[i]
        r.out_depth, r.out_flux4, r.out_velocity2 = _ptr(out["depth"]), _ptr(out["flux"]), \
            _ptr(out["velocity"])
        ticket = C.c_uint64()
        check(lib().tws_server_submit(self._h, C.byref(r), C.byref(ticket)))
        self._keep[ticket.value] = out
        return ticket.value, out

    def release(self, ticket):
        """Forgets a finished request: the server's record and the output
        arrays kept for it (tws_server_release)."""
        check(lib().tws_server_release(self._h, ticket))
        self._keep.pop(ticket, None)

    def poll(self, ticket) -> RequestStatus:
        st = RequestStatus()
        check(lib().tws_server_poll(self._h, ticket, C.byref(st)))
        return st

    def wait(self, ticket, timeout=120.0) -> RequestStatus:
        st = RequestStatus()
        check(lib().tws_server_wait(self._h, ticket, float(timeout), C.byref(st)))
        return st

    def cancel(self, ticket):
        check(lib().tws_server_cancel(self._h, ticket))

    def drain(self):
        check(lib().tws_server_drain(self._h))

    def now(self) -> float:
        return lib().tws_server_now(self._h)

    def stats(self) -> dict:
        st = ServerStats()
        check(lib().tws_server_stats_get(self._h, C.byref(st)))
        return {name: getattr(st, name) for name, _ in st._fields_}
