"""C-ABI checks that need no GPU: the library loads, exports every function
include/tws.h declares, and its host-only entry points behave."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
from bitwise import same_bits
import pytest

ROOT = Path(__file__).resolve().parent.parent


def declared_functions():
    names = set()
    for h in sorted((ROOT / "include").glob("*.h")):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names |= set(re.findall(r"\b(tws_\w+)\s*\(", text))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_functions()
    assert "tws_step" in names and "tws_create" in names and len(names) >= 20


def test_library_exports_every_declared_symbol(tws):
    lib = tws.lib()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"declared in include/*.h but not exported: {missing}"


def test_binding_covers_every_declared_symbol(tws):
    assert set(declared_functions()) <= set(tws.PROTOTYPES)


def test_abi_version(tws):
    assert tws.lib().tws_abi_version() == 1


def test_params_match_oracle(tws, oracle):
    p = tws.params_reference(60.0, 0.98, 10.0, 1024.0, 1024)
    assert (p.flow_friction, p.water_acceleration, p.cell_area_inv_timescaled) == \
        oracle.params_reference(60.0, 0.98, 10.0, 1024.0, 1024)
    p = tws.params_reference(30.0, 0.96, 30.0, 1024.0, 512)
    assert (p.flow_friction, p.water_acceleration, p.cell_area_inv_timescaled) == \
        oracle.params_reference(30.0, 0.96, 30.0, 1024.0, 512)
    q = tws.params_physical(0.02, 9.81, 1.0, 1.0, 0.995)
    assert (q.flow_friction, q.water_acceleration, q.cell_area_inv_timescaled) == \
        oracle.params_physical(0.02, 9.81, 1.0, 1.0, 0.995)
    assert abs(q.water_acceleration - 0.1962) < 1e-7
    assert abs(q.cell_area_inv_timescaled - 0.02) < 1e-9
    assert abs(q.flow_friction - 0.995 ** 0.02) < 1e-7


def test_invalid_arguments_are_rejected(tws):
    with pytest.raises(tws.TwsError) as e:
        tws.params_physical(dt=-1.0)
    assert e.value.code == -1 and "invalid" in str(e.value)
    with pytest.raises(tws.TwsError):
        tws.params_reference(boundary=7)
    p = tws.Params()
    assert tws.lib().tws_params_physical(0.02, 9.81, 1.0, 1.0, 0.995, 0, None) == -1
    assert tws.lib().tws_step(None, 1) == -1
    assert tws.lib().tws_destroy(None) == 0
    del p
    # F, K, C must be finite and >= 0 (checked before any device call)
    import ctypes as C
    h = C.c_void_p()
    for F, K, Cc in ((0.99, -0.1, 0.02), (-0.5, 0.2, 0.02), (0.99, 0.2, -0.02),
                     (float("nan"), 0.2, 0.02), (0.99, float("inf"), 0.02)):
        bad = tws.params_raw(F, K, Cc)
        assert tws.lib().tws_create(64, 64, C.byref(bad), None, C.byref(h)) == -1
        assert not h.value


def test_fbm_terrain_is_deterministic(tws):
    a = tws.generate_fbm_terrain(96, 64, seed=1)
    b = tws.generate_fbm_terrain(96, 64, seed=1)
    c = tws.generate_fbm_terrain(96, 64, seed=2)
    assert same_bits(a, b) and not same_bits(a, c)
    assert a.min() >= 0.0 and a.max() <= 100.0
    assert a.std() > 2.0  # has relief
    # period-wrapped lattice: the field continues smoothly across the seam
    big = tws.generate_fbm_terrain(256, 256, seed=3)
    assert np.abs(big[:, 0] - big[:, -1]).max() < np.abs(big[:, 0] - big[:, 128]).max()
    # multithreaded generation matches row-by-row generation
    rows = np.stack([tws.generate_fbm_terrain(256, 256, seed=3)[y] for y in (0, 100, 255)])
    assert same_bits(rows, big[[0, 100, 255]])


def test_tune_record_validates_and_forget_clears(tws):
    lib = tws.lib()
    assert lib.tws_tune_record(0, 64, tws.KERNEL_TILE_128x64, 4) == -1
    assert lib.tws_tune_record(64, 64, 99, 4) == -1
    assert lib.tws_tune_record(64, 64, tws.KERNEL_TILE_128x64, 7) == -1
    assert b"tws_tune_record" in lib.tws_last_error()
    assert lib.tws_tune_record(64, 64, tws.KERNEL_STREAM_128, 7) == -1
    assert lib.tws_tune_record(64, 64, tws.KERNEL_TILE_128x64, 4) == 0
    assert lib.tws_tune_record(4096, 4096, tws.KERNEL_STREAM_128, 6) == 0
    assert lib.tws_tune_forget() == 0


def test_closed_walls_bound_k(tws):
    """Closed walls need the wall ghost's head difference times K to stay
    finite: K above 1e8 is refused with closed walls, accepted otherwise."""
    import ctypes as C
    h = C.c_void_p()
    bad = tws.params_raw(0.99, 2e8, 0.02, tws.BOUNDARY_CLOSED)
    assert tws.lib().tws_create(64, 64, C.byref(bad), None, C.byref(h)) == -1
    assert not h.value
    for bnd in (tws.BOUNDARY_ZERO, tws.BOUNDARY_REFERENCE_DISPATCH):
        p = tws.params_physical(0.02, 9.81, 1.0, 1.0, 0.995, boundary=bnd)
        assert p.boundary == bnd
    with pytest.raises(tws.TwsError):
        tws.params_physical(0.02, 9.81, 1.0, 1.0, 0.995, boundary=3)


def test_stream_copy_shapes(tws):
    assert tws.stream_copy_shapes() >= 4
    assert tws.lib().tws_stream_copy_bench(1 << 20, 1, tws.stream_copy_shapes(), None) == -1
