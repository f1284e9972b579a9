"""ctypes binding of libtws.so (the C ABI declared in include/tws.h).

The library is loaded from this package directory.  There is no fallback: if
libtws.so is missing every entry point raises, naming the build command.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

_PKG = Path(__file__).resolve().parent
LIB_PATH = Path(os.environ.get("TWS_LIBRARY", _PKG / "libtws.so"))

TWS_OK = 0
TWS_ERR_INVALID = -1
TWS_ERR_NOMEM = -2
TWS_ERR_DEVICE = -3
TWS_ERR_NONFINITE = -4
TWS_ERR_STATE = -5
REQ_QUEUED, REQ_RUNNING, REQ_DONE, REQ_FAILED, REQ_CANCELLED = 0, 1, 2, 3, 4
TWS_ERR_CAPACITY = -6

BOUNDARY_CLOSED = 0
BOUNDARY_ZERO = 1
BOUNDARY_REFERENCE_DISPATCH = 2

KERNEL_AUTO = 0
KERNEL_TWOPASS = 1
KERNEL_TILE_128x128 = 2
KERNEL_TILE_128x64 = 3
KERNEL_TILE_256x64 = 4
KERNEL_TILE_64x64 = 5
KERNEL_TILE_64x64_16W = 6
KERNEL_TILE_64x48_16W = 7
KERNEL_STREAM_128 = 8
BATCH_MAX_CLASSES = 4
TILE_KERNELS = (KERNEL_TILE_128x128, KERNEL_TILE_128x64, KERNEL_TILE_256x64, KERNEL_TILE_64x64,
                KERNEL_TILE_64x64_16W, KERNEL_TILE_64x48_16W)


class TwsError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"[{code}] {message}")
        self.code = code


class Params(C.Structure):
    _fields_ = [
        ("flow_friction", C.c_float),
        ("water_acceleration", C.c_float),
        ("cell_area_inv_timescaled", C.c_float),
        ("boundary", C.c_int32),
        ("ghost_alpha", C.c_float),
    ]

    def as_tuple(self):
        return (self.flow_friction, self.water_acceleration, self.cell_area_inv_timescaled,
                self.boundary, self.ghost_alpha)


class DevicePlanes(C.Structure):
    _fields_ = [
        ("terrain", C.c_void_p),
        ("depth", C.c_void_p),
        ("flux", C.c_void_p * 4),
        ("pitch", C.c_int64),
        ("width", C.c_int32),
        ("height", C.c_int32),
    ]


class BatchStats(C.Structure):
    _fields_ = [
        ("grids", C.c_int32),
        ("max_grids", C.c_int32),
        ("cells", C.c_int64),
        ("pool_bytes", C.c_int64),
        ("pool_used", C.c_int64),
        ("pool_largest_free", C.c_int64),
        ("graph_captures", C.c_int64),
        ("graph_replays", C.c_int64),
    ]


class TuneEntry(C.Structure):
    _fields_ = [("kernel", C.c_int32), ("steps_per_launch", C.c_int32),
                ("cell_updates_per_s", C.c_double)]


class ServerConfig(C.Structure):
    _fields_ = [
        ("pool_bytes", C.c_int64),
        ("max_resident", C.c_int32),
        ("chunk_steps", C.c_int32),
        ("kernel", C.c_int32),
        ("steps_per_launch", C.c_int32),
        ("max_queue", C.c_int32),
    ]


class Request(C.Structure):
    _fields_ = [
        ("width", C.c_int32), ("height", C.c_int32), ("steps", C.c_int32),
        ("params", Params),
        ("terrain", C.c_void_p), ("depth", C.c_void_p), ("flux4", C.c_void_p),
        ("rain_on", C.c_int32), ("rain", C.c_float * 4),
        ("out_depth", C.c_void_p), ("out_flux4", C.c_void_p), ("out_velocity2", C.c_void_p),
    ]


class RequestStatus(C.Structure):
    _fields_ = [
        ("state", C.c_int32), ("error", C.c_int32), ("steps_done", C.c_int32), ("pad_", C.c_int32),
        ("t_submit", C.c_double), ("t_admit", C.c_double), ("t_done", C.c_double),
    ]


class ServerStats(C.Structure):
    _fields_ = [
        ("submitted", C.c_int64), ("completed", C.c_int64), ("failed", C.c_int64),
        ("cancelled", C.c_int64), ("queued", C.c_int64), ("resident", C.c_int64),
        ("runs", C.c_int64), ("cell_updates", C.c_int64), ("busy_seconds", C.c_double),
        ("pool_used", C.c_int64), ("pool_bytes", C.c_int64),
    ]


_fp = C.POINTER(C.c_float)
_vp = C.c_void_p
_i32 = C.c_int32

# name -> (restype, argtypes); every function declared in include/tws.h.
PROTOTYPES = {
    "tws_params_physical": (C.c_int, [C.c_float, C.c_float, C.c_float, C.c_float, C.c_float, _i32, C.POINTER(Params)]),
    "tws_params_reference": (C.c_int, [C.c_float, C.c_float, C.c_float, C.c_float, C.c_uint32, _i32, C.POINTER(Params)]),
    "tws_create": (C.c_int, [_i32, _i32, C.POINTER(Params), _vp, C.POINTER(_vp)]),
    "tws_destroy": (C.c_int, [_vp]),
    "tws_set_params": (C.c_int, [_vp, C.POINTER(Params)]),
    "tws_get_params": (C.c_int, [_vp, C.POINTER(Params)]),
    "tws_upload": (C.c_int, [_vp, _vp, _vp, _vp]),
    "tws_download": (C.c_int, [_vp, _vp, _vp, _vp]),
    "tws_upload_reference_layout": (C.c_int, [_vp, _vp, _vp]),
    "tws_download_reference_layout": (C.c_int, [_vp, _vp, C.c_float, _vp, _vp]),
    "tws_apply_brush": (C.c_int, [_vp, C.c_float, C.c_float, C.c_float, C.c_float]),
    "tws_set_rain": (C.c_int, [_vp, C.c_float, C.c_float, C.c_float, C.c_float]),
    "tws_step": (C.c_int, [_vp, _i32]),
    "tws_step_profiled": (C.c_int, [_vp, _i32, _fp, _i32, C.POINTER(_i32)]),
    "tws_perform_simulation_step": (C.c_int, [_vp, C.c_double, C.c_double, C.POINTER(_i32)]),
    "tws_mip_chain_size": (C.c_int, [_i32, _i32, C.POINTER(_i32), C.POINTER(C.c_size_t)]),
    "tws_generate_mips": (C.c_int, [_vp, C.c_float, _vp, C.c_size_t]),
    "tws_download_mips": (C.c_int, [_vp, C.c_float, _vp, C.c_size_t]),
    "tws_set_frame_mips": (C.c_int, [_vp, _vp, C.c_size_t, C.c_float]),
    "tws_volume": (C.c_int, [_vp, C.POINTER(C.c_double)]),
    "tws_check_finite": (C.c_int, [_vp]),
    "tws_sync": (C.c_int, [_vp]),
    "tws_get_device_planes": (C.c_int, [_vp, C.POINTER(DevicePlanes)]),
    "tws_set_launch": (C.c_int, [_vp, _i32, _i32]),
    "tws_get_launch": (C.c_int, [_vp, C.POINTER(_i32), C.POINTER(_i32)]),
    "tws_set_tile_order": (C.c_int, [_vp, _i32]),
    "tws_generate_fbm_terrain": (C.c_int, [_i32, _i32, C.c_uint64, _i32, _i32, C.c_float, C.c_float, _vp]),
    "tws_generate_reference_scene": (C.c_int, [_i32, C.c_float, C.c_uint32, _i32, _vp, _vp]),
    "tws_init_terrain_fbm": (C.c_int, [_vp, C.c_uint64, _i32, _i32, C.c_float, C.c_float]),
    "tws_init_water": (C.c_int, [_vp, C.c_float, _i32, _i32, _i32, _i32, C.c_float]),
    "tws_batch_init_terrain_fbm": (C.c_int, [_vp, _i32, C.c_uint64, _i32, _i32, C.c_float, C.c_float]),
    "tws_batch_init_water": (C.c_int, [_vp, _i32, C.c_float, _i32, _i32, _i32, _i32, C.c_float]),
    "tws_init_reference_scene": (C.c_int, [_vp, C.c_float, C.c_uint32, _i32]),
    "tws_batch_init_reference_scene": (C.c_int, [_vp, _i32, C.c_float, C.c_uint32, _i32]),
    "tws_debug_lane_shift_probe": (C.c_int, [C.POINTER(_i32)]),
    "tws_float_to_half_rtz": (C.c_int, [_vp, _vp, C.c_int64]),
    "tws_stream_copy_bench": (C.c_int, [C.c_int64, _i32, _i32, C.POINTER(C.c_double)]),
    "tws_stream_copy_shapes": (C.c_int, []),
    "tws_grid_footprint": (C.c_int64, [_i32, _i32]),
    "tws_batch_create": (C.c_int, [_i32, C.c_int64, _vp, C.POINTER(_vp)]),
    "tws_batch_destroy": (C.c_int, [_vp]),
    "tws_batch_add": (C.c_int, [_vp, _i32, _i32, C.POINTER(Params), C.POINTER(_i32)]),
    "tws_batch_remove": (C.c_int, [_vp, _i32]),
    "tws_batch_upload": (C.c_int, [_vp, _i32, _vp, _vp, _vp]),
    "tws_batch_download": (C.c_int, [_vp, _i32, _vp, _vp, _vp]),
    "tws_batch_set_params": (C.c_int, [_vp, _i32, C.POINTER(Params)]),
    "tws_batch_set_rain": (C.c_int, [_vp, _i32, C.c_float, C.c_float, C.c_float, C.c_float]),
    "tws_batch_volume": (C.c_int, [_vp, _i32, C.POINTER(C.c_double)]),
    "tws_batch_check_finite": (C.c_int, [_vp, _i32]),
    "tws_batch_set_steps": (C.c_int, [_vp, _i32, _i32]),
    "tws_batch_run": (C.c_int, [_vp]),
    "tws_batch_step": (C.c_int, [_vp, _i32]),
    "tws_batch_set_launch": (C.c_int, [_vp, _i32, _i32]),
    "tws_batch_get_launch": (C.c_int, [_vp, C.POINTER(_i32), C.POINTER(_i32)]),
    "tws_batch_set_launch_classes": (C.c_int, [_vp, _i32, _vp, _vp, _vp]),
    "tws_batch_get_launch_classes": (C.c_int, [_vp, _i32, _vp, _vp, _vp, C.POINTER(_i32)]),
    "tws_batch_use_graphs": (C.c_int, [_vp, _i32]),
    "tws_batch_stats_get": (C.c_int, [_vp, C.POINTER(BatchStats)]),
    "tws_batch_sync": (C.c_int, [_vp]),
    "tws_autotune": (C.c_int, [_vp, _vp, _i32, _i32, C.POINTER(TuneEntry), C.POINTER(TuneEntry), _i32, C.POINTER(_i32)]),
    "tws_tune_forget": (C.c_int, []),
    "tws_volume_noalloc": (C.c_int, [_vp, _i32, _i32, C.c_int64, _vp, C.POINTER(C.c_double)]),
    "tws_tune_record": (C.c_int, [_i32, _i32, _i32, _i32]),
    "tws_server_create": (C.c_int, [C.POINTER(ServerConfig), C.POINTER(_vp)]),
    "tws_server_submit": (C.c_int, [_vp, C.POINTER(Request), C.POINTER(C.c_uint64)]),
    "tws_server_poll": (C.c_int, [_vp, C.c_uint64, C.POINTER(RequestStatus)]),
    "tws_server_wait": (C.c_int, [_vp, C.c_uint64, C.c_double, C.POINTER(RequestStatus)]),
    "tws_server_cancel": (C.c_int, [_vp, C.c_uint64]),
    "tws_server_drain": (C.c_int, [_vp]),
    "tws_server_release": (C.c_int, [_vp, C.c_uint64]),
    "tws_server_stats_get": (C.c_int, [_vp, C.POINTER(ServerStats)]),
    "tws_server_now": (C.c_double, [_vp]),
    "tws_server_destroy": (C.c_int, [_vp]),
    "tws_last_error": (C.c_char_p, []),
    "tws_abi_version": (C.c_int, []),
    "tws_device_count": (C.c_int, []),
}

_lib = None


def lib() -> C.CDLL:
    """Loads libtws.so once; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m terrainwatersim_amd.build` "
                "(hipcc, gfx950). There is no CPU fallback.")
        handle = C.CDLL(str(LIB_PATH))
        for name, (res, args) in PROTOTYPES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def check(code: int) -> int:
    if code != TWS_OK:
        msg = lib().tws_last_error()
        raise TwsError(code, msg.decode() if msg else "unknown error")
    return code
