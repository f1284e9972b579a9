"""The 16384^2 dam-break at full size: a 1024^2 window at the centre after
256 steps against the FP32 restatement (bitwise) and the FP64 restatement
(error reported), both run on a 2052^2 sub-domain that holds the window's
whole dependency cone (depth reaches 2 cells per step), plus volume
conservation over 1000 steps of the full grid."""
import numpy as np
from bitwise import same_bits
import pytest

pytestmark = pytest.mark.gpu


def test_large_dam_break_window(gpu, oracle):
    from terrainwatersim_amd.workloads import PARAMS, large_state
    N, window, t = 16384, 1024, 256
    p = gpu.params_physical(**PARAMS)
    b, d = large_state(N, 5)
    w0 = (N - window) // 2
    with gpu.Simulation(N, N, p) as sim:
        sim.upload(b, d)
        v0 = sim.volume()
        sim.step(t)
        win = sim.download(flux=False)["depth"][w0:w0 + window, w0:w0 + window].copy()
        sim.step(1000 - t)
        v1 = sim.volume()
        assert sim.check_finite()
    assert abs(v1 - v0) / v0 < 1e-6, (v0, v1)
    pad = 2 * t + 2
    s0, s1 = w0 - pad, w0 + window + pad
    bs = np.ascontiguousarray(b[s0:s1, s0:s1])
    ds = np.ascontiguousarray(d[s0:s1, s0:s1])
    del b, d
    oracle.set_threads(16)
    F, K, Cc = p.flow_friction, p.water_acceleration, p.cell_area_inv_timescaled
    d32, _, _ = oracle.step(bs, ds, None, F, K, Cc, n=t)
    assert same_bits(win, d32[pad:pad + window, pad:pad + window])
    d64, _, _ = oracle.step(bs, ds, None, F, K, Cc, n=t, dtype=np.float64)
    r64 = d64[pad:pad + window, pad:pad + window]
    m = float(r64.max())
    err = np.abs(win - r64)
    print(f"16384^2 dam-break window {window}^2 after {t} steps: FP32 vs FP64 max abs "
          f"{err.max():.3e} m, max rel (floor 1e-3*max) "
          f"{(err / np.maximum(np.abs(r64), 1e-3 * m)).max():.3e}; volume drift over 1000 "
          f"steps {abs(v1 - v0) / v0:.3e}")
    assert err.max() < 1e-3 * m
