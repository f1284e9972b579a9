"""Terrain mip chain on the GPU (through the C ABI) against the restatement and
llvmpipe's glGenerateMipmap fixtures.

Bar: bit-identical to the restatement for every shape, and to llvmpipe where
every level halves (power-of-two grids); non-power-of-two fixtures within
llvmpipe's sample-position rounding (tests/test_mips.py).  The frame driver
regenerates the chain into the caller's device buffer only after frames that
ran steps (Terrain.cpp:270-276).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from test_mips import NPOT, NPOT_TOL, POT, fixture, rgba

pytestmark = pytest.mark.gpu
F32 = np.float32


class DeviceBuffer:
    """Device memory from the HIP runtime libtws.so itself uses."""

    def __init__(self, nfloats):
        self.hip = C.CDLL("libamdhip64.so.7")
        self.n = int(nfloats)
        self.ptr = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.ptr), C.c_size_t(4 * max(self.n, 1))) == 0

    def fill(self, host):
        a = np.ascontiguousarray(host, F32)
        assert a.size == self.n
        assert self.hip.hipMemcpy(self.ptr, a.ctypes.data_as(C.c_void_p), C.c_size_t(4 * self.n),
                                  1) == 0  # hipMemcpyHostToDevice
        assert self.hip.hipDeviceSynchronize() == 0

    def read(self):
        out = np.empty(self.n, F32)
        assert self.hip.hipDeviceSynchronize() == 0
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, C.c_size_t(4 * self.n),
                                  2) == 0  # hipMemcpyDeviceToHost
        return out

    def free(self):
        if self.ptr:
            self.hip.hipFree(self.ptr)
            self.ptr = C.c_void_p()


def flat(levels):
    return np.concatenate([a.ravel() for a in levels]) if levels else np.zeros(0, F32)


def state(W, H, seed):
    rng = np.random.default_rng(seed)
    b = (rng.random((H, W)) * 40.0).astype(F32)
    d = (rng.random((H, W)) * 2.0 * (rng.random((H, W)) < 0.6)).astype(F32)
    return b, d


@pytest.mark.parametrize("W,H", [(256, 256), (128, 32), (1, 16), (16, 1), (100, 60), (37, 91),
                                 (1000, 1000), (4096, 4096)])
def test_hip_mips_match_restatement_bitwise(gpu, oracle, W, H):
    print(f"mips {W}x{H}", flush=True)
    b, d = state(W, H, W * 7 + H)
    with gpu.Simulation(W, H) as s:
        s.upload(b, d)
        ours = s.mips(0.3)
    want = oracle.mips(rgba(b, d))
    assert len(ours) == len(want) == oracle.mip_levels(W, H) - 1
    for l, (a, w) in enumerate(zip(ours, want), 1):
        assert a.shape == w.shape
        assert np.array_equal(a.view(np.uint32), w.view(np.uint32)), f"level {l}"


@pytest.mark.parametrize("name", POT + NPOT)
def test_hip_mips_match_llvmpipe(gpu, name):
    b, d, gold = fixture(name)
    H, W = b.shape
    with gpu.Simulation(W, H) as s:
        s.upload(b, d)
        ours = s.mips(0.3)
    span = max(np.ptp(b), np.ptp(d))
    for l, (a, g) in enumerate(zip(ours, gold), 1):
        if name in POT:
            assert np.array_equal(a.view(np.uint32), g.view(np.uint32)), f"level {l}"
        else:
            assert np.max(np.abs(a - g)) <= NPOT_TOL * span, f"level {l}"


def test_generate_into_device_buffer(gpu):
    W, H = 300, 200
    b, d = state(W, H, 11)
    levels, floats = gpu.mip_chain_size(W, H)
    buf = DeviceBuffer(floats)
    try:
        with gpu.Simulation(W, H) as s:
            s.upload(b, d)
            s.generate_mips(buf.ptr.value, floats, 0.5)
            s.sync()
            assert np.array_equal(buf.read(), flat(s.mips(0.5)))
            with pytest.raises(gpu.TwsError):
                s.generate_mips(buf.ptr.value, floats - 4, 0.5)
            with pytest.raises(gpu.TwsError):
                s.generate_mips(buf.ptr.value + 4, floats - 4, 0.5)  # not 16-byte aligned
    finally:
        buf.free()


def test_frame_driver_regenerates_after_frames_with_steps(gpu, oracle):
    W = H = 64
    b, d = state(W, H, 12)
    p = gpu.params_physical()
    levels, floats = gpu.mip_chain_size(W, H)
    buf = DeviceBuffer(floats)
    sentinel = np.full(floats, -1.0, F32)
    try:
        with gpu.Simulation(W, H, p) as s:
            s.upload(b, d)
            s.set_frame_mips(buf.ptr.value, floats, 0.3)
            buf.fill(sentinel)
            assert s.perform_simulation_step(0.004, 0.01) == 0   # no step: chain untouched
            s.sync()
            assert np.array_equal(buf.read(), sentinel)
            assert s.perform_simulation_step(0.03, 0.01) == 3    # 0.034 s -> 3 steps
            s.sync()
            got = buf.read()
            st = s.download(flux=False)["depth"]
            assert np.array_equal(got, flat(oracle.mips(rgba(b, st))))
            assert np.array_equal(got, flat(s.mips(0.3)))
            s.set_frame_mips(None)
            buf.fill(sentinel)
            assert s.perform_simulation_step(0.02, 0.01) == 2
            s.sync()
            assert np.array_equal(buf.read(), sentinel)
    finally:
        buf.free()
