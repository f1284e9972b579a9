"""tws_volume_noalloc (allocates no device memory) against tws_volume, through
the C ABI: bit-identical on two calls in a row, the second on a smaller grid,
and on a plane with padded rows."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(tws):
    fn = tws.lib().tws_volume_noalloc
    fn.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int64, C.c_void_p,
                   C.POINTER(C.c_double)]
    fn.restype = C.c_int
    return fn


def _volumes(tws, W, H, seed):
    rng = np.random.default_rng(seed)
    b = (rng.random((H, W)) * 5.0).astype(np.float32)
    d = (rng.random((H, W)) * 3.0 * (rng.random((H, W)) < 0.8)).astype(np.float32)
    p = tws.params_physical(0.02, 9.81, 1.0, 1.0, 0.995)
    with tws.Simulation(W, H, p) as s:
        s.upload(b, d)
        s.step(3)
        s.sync()
        ref = s.volume()
        planes = s.device_planes()
        out = C.c_double(float("nan"))
        # the simulation's own stream is not exposed; NULL (the default
        # stream) after the sync above sees the finished state
        rc = _noalloc(tws)(C.c_void_p(planes.depth), W, H, planes.pitch, None, C.byref(out))
        assert rc == 0, tws.lib().tws_last_error()
        host = s.download(flux=False)["depth"]
    return ref, out.value, float(np.sum(host.astype(np.float64)))


def test_volume_noalloc_matches_tws_volume(gpu):
    print("no-alloc volume: 1000 x 700, then 130 x 97", flush=True)
    ref1, got1, host1 = _volumes(gpu, 1000, 700, 1)
    assert got1 == ref1, (got1, ref1)
    assert abs(got1 - host1) <= 1e-9 * host1
    # second call, smaller grid: the scratch holds the first call's partials
    ref2, got2, host2 = _volumes(gpu, 130, 97, 2)
    assert got2 == ref2, (got2, ref2)
    assert abs(got2 - host2) <= 1e-9 * host2


def test_volume_noalloc_rejects_bad_planes(gpu):
    fn = _noalloc(gpu)
    out = C.c_double()
    dummy = C.c_void_p(16)
    for W, H, pitch in ((0, 10, 64), (10, 0, 64), (100, 10, 64), (1 << 21, 1, 1 << 21)):
        assert fn(dummy, W, H, pitch, None, C.byref(out)) == -1  # TWS_ERR_INVALID
        assert b"tws_volume_noalloc" in gpu.lib().tws_last_error()
    assert fn(None, 10, 10, 64, None, C.byref(out)) == -1
