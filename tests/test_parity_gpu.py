"""Parity of the HIP step (through the C ABI) with the CPU restatement.

Bar: bit-identical d, f and v for the FP32 path (the library and the oracle
are both built without multiply-add contraction and keep the shaders'
operation order).  Full-size grids are checked through size-independent
properties: volume conservation, non-negativity, blocked == unblocked, and a
bitwise comparison of the first steps against the oracle.
"""
import numpy as np
from bitwise import same_bits
import pytest

pytestmark = pytest.mark.gpu

F32 = np.float32
CFG_PARAMS = dict(dt=0.02, gravity=9.81, cell_size=1.0, pipe_area=1.0, damping=0.995)


def random_state(seed, H, W, wet=0.7):
    rng = np.random.default_rng(seed)
    b = (rng.random((H, W)) * 5.0).astype(F32)
    d = (rng.random((H, W)) * 2.0 * (rng.random((H, W)) < wet)).astype(F32)
    f = (rng.random((H, W, 4)) * 0.5).astype(F32)
    return b, d, f


def run_gpu(tws, b, d, f, n, kernel, k, params, rain=None, velocity=True):
    H, W = b.shape
    with tws.Simulation(W, H, params) as s:
        s.set_launch(kernel, k)
        s.upload(b, d, f)
        if rain is not None:
            s.set_rain(*rain)
        s.step(n)
        out = s.download(depth=True, flux=True, velocity=velocity)
        ok = s.check_finite()
    assert ok
    return out


def p3(p):
    return p.flow_friction, p.water_acceleration, p.cell_area_inv_timescaled


def assert_same(out, d, f, v=None):
    assert same_bits(out["depth"], d), f"depth differs at {np.argwhere(out['depth'] != d)[:5]}"
    assert same_bits(out["flux"], f), "flux differs"
    if v is not None:
        assert same_bits(out["velocity"], v), "velocity differs"


def test_lane_shift_layout(gpu):
    # lane i holds i + 1; a lane with no source lane reads 0
    out = gpu.lane_shift_probe()
    assert out[:64].tolist() == list(range(64))
    assert out[64:].tolist() == list(range(2, 65)) + [0]


KERNELS = [(2, 1), (2, 2), (2, 4), (2, 8), (3, 1), (3, 4), (3, 5), (3, 8), (4, 1), (4, 2), (4, 4),
           (4, 5), (4, 8), (5, 1), (5, 4), (5, 5), (5, 6), (5, 8), (6, 1), (6, 3), (6, 5), (6, 8),
           (7, 1), (7, 2), (7, 4), (7, 6), (7, 8), (8, 2), (8, 3), (8, 4), (8, 5), (8, 6),
           (8, 8)]


@pytest.mark.parametrize("kernel,k", KERNELS)
def test_tile_kernel_bitwise(gpu, oracle, kernel, k):
    b, d, f = random_state(10 + kernel * 10 + k, 157, 301)
    p = gpu.params_physical(**CFG_PARAMS)
    rain = (100.0, 60.0, 1e-3, 256.0)
    n = 11
    out = run_gpu(gpu, b, d, f, n, kernel, k, p, rain=rain)
    de, fe, ve = oracle.step(b, d, f, *p3(p), n=n, rain=rain, velocity=True)
    assert_same(out, de, fe, ve)


RAIN_DISCS = [(100.0, 60.0, 1e-3, 256.0),     # inside
              (0.5, 199.5, 2e-3, 900.0),       # centre on a corner, most of the disc outside
              (111.9, 47.9, 1e-2, 0.5),        # under one cell, next to tile seams
              (-30.0, 80.0, 1e-3, 1024.0),     # centre outside, 2 columns inside
              (150.0, 100.0, 1e-4, 1e30),      # covers the whole grid
              (5000.0, -700.0, 1e-3, 400.0),   # far outside: no cell
              (64.25, 56.75, 5e-3, 3.0)]       # a few cells at a wave's first rows


@pytest.mark.parametrize("kernel,k", [(3, 4), (3, 5), (5, 8), (2, 3), (4, 6), (8, 5), (8, 3)])
def test_rain_disc_positions(gpu, oracle, kernel, k):
    """The step kernels skip the brush for waves that the disc cannot reach; a
    disc at every kind of position (under a single cell, across tile seams,
    over a corner, partly or wholly outside, covering everything) gives the
    restatement's depth and flux bit for bit."""
    b, d, f = random_state(77 + kernel * 10 + k, 200, 300)
    p = gpu.params_physical(**CFG_PARAMS)
    for rain in RAIN_DISCS:
        out = run_gpu(gpu, b, d, f, 9, kernel, k, p, rain=rain, velocity=False)
        de, fe, _ = oracle.step(b, d, f, *p3(p), n=9, rain=rain)
        assert same_bits(out["depth"], de), rain
        assert same_bits(out["flux"], fe), rain


@pytest.mark.parametrize("kernel,k", [(1, 1), (3, 5), (5, 8), (7, 8), (8, 6), (8, 4)])
def test_negative_zero_depth_bitwise(gpu, oracle, kernel, k):
    """Cells uploaded with a depth of -0: on a slope their outflow exceeds
    their depth, so the flux is scaled by -0 / total = -0 (flowUpdate.comp:
    58-59), which signs the scaled fluxes.  Where such a cell receives only
    -0 inflows its new depth is max(0, -0), which GLSL's max makes +0
    (flowApply.comp:41).  One step (each kernel family's one-step form) and k
    steps (its k-step form), compared by bit pattern."""
    b, d, f = random_state(90 + kernel * 10 + k, 130, 250)
    rng = np.random.default_rng(kernel * 100 + k)
    d[rng.random(d.shape) < 0.3] = F32(-0.0)
    p = gpu.params_physical(**CFG_PARAMS)
    for n in sorted({1, k}):
        out = run_gpu(gpu, b, d, f, n, kernel, k, p)
        de, fe, ve = oracle.step(b, d, f, *p3(p), n=n, velocity=True)
        if n == 1:
            assert np.signbit(fe).any()  # the case the test is for
        assert_same(out, de, fe, ve)


def test_twopass_bitwise(gpu, oracle):
    b, d, f = random_state(3, 90, 70)
    p = gpu.params_physical(**CFG_PARAMS)
    out = run_gpu(gpu, b, d, f, 9, gpu.KERNEL_TWOPASS, 1, p, rain=(30.0, 20.0, 1e-2, 100.0))
    de, fe, ve = oracle.step(b, d, f, *p3(p), n=9, rain=(30.0, 20.0, 1e-2, 100.0), velocity=True)
    assert_same(out, de, fe, ve)


@pytest.mark.parametrize("kernel,k", [(1, 1), (2, 4), (3, 2), (4, 8), (3, 5), (5, 4), (6, 8),
                                      (7, 6), (7, 8), (8, 5)])
def test_zero_boundary_bitwise(gpu, oracle, kernel, k):
    b, d, f = random_state(21, 130, 259)
    p = gpu.params_physical(**CFG_PARAMS, boundary=gpu.BOUNDARY_ZERO)
    out = run_gpu(gpu, b, d, f, 13, kernel, k, p)
    de, fe, ve = oracle.step(b, d, f, *p3(p), n=13, boundary=oracle.BOUNDARY_ZERO, velocity=True)
    assert_same(out, de, fe, ve)


@pytest.mark.parametrize("shape", [(1, 1), (1, 7), (5, 1), (2, 3), (64, 64), (129, 257),
                                   (1, 5000), (3000, 2), (45, 109)])
def test_ragged_and_tiny_grids(gpu, oracle, shape):
    H, W = shape
    b, d, f = random_state(H * 1000 + W, H, W)
    p = gpu.params_physical(**CFG_PARAMS)
    for kernel, k in ((2, 4), (4, 8), (1, 1), (3, 5), (5, 6), (6, 8), (7, 8), (8, 5), (8, 8)):
        out = run_gpu(gpu, b, d, f, 6, kernel, k, p)
        de, fe, ve = oracle.step(b, d, f, *p3(p), n=6, velocity=True)
        assert_same(out, de, fe, ve)


def config1_state(tws):
    b = tws.generate_fbm_terrain(256, 256, seed=1)
    d = np.zeros((256, 256), F32)
    return b, d


def test_config1_rain_1000_steps(gpu, oracle):
    """256^2 fBm terrain, rain disc r=16 adding 1e-3 m/step at the centre,
    closed walls, 1000 steps: bitwise vs the FP32 restatement, volume equal to
    the FP64 sum of deposits to 1e-6, FP32 vs FP64 error reported."""
    b, d = config1_state(gpu)
    p = gpu.params_physical(**CFG_PARAMS)
    rain = (128.0, 128.0, 1e-3, 256.0)
    with gpu.Simulation(256, 256, p) as s:
        s.upload(b, d)
        s.set_rain(*rain)
        s.step(1000)
        out = s.download(velocity=True)
        vol = s.volume()
    de, fe, ve = oracle.step(b, d, None, *p3(p), n=1000, rain=rain, velocity=True)
    assert_same(out, de, fe, ve)
    dep = float(oracle.brush(np.zeros((256, 256)), *rain, dtype=np.float64).sum()) * 1000
    assert abs(vol - dep) / dep < 1e-6
    d64, _, _ = oracle.step(b, d, None, *p3(p), n=1000, rain=rain, dtype=np.float64)
    m = float(d64.max())
    rel = np.abs(out["depth"] - d64) / np.maximum(np.abs(d64), 1e-3 * m)
    print(f"config1 FP32 vs FP64 after 1000 steps: max rel (floor 1e-3*max) {rel.max():.3e}, "
          f"max abs {np.abs(out['depth'] - d64).max():.3e} m, max depth {m:.4f} m")
    # FP32 rounding, not a parity bar (the bar is the bitwise check above):
    # measured 2.2e-3 relative (floor 1e-3 * max depth), 5.1e-5 m absolute.
    assert np.abs(out["depth"] - d64).max() < 1e-4 * m


def test_full_size_4096_properties(gpu, oracle):
    """4096^2, uniform 0.5 m sheet on fBm terrain (the HBM-bound workload):
    first 2 steps bitwise vs the oracle; 1000 steps conserve volume and keep
    d >= 0; after 200 steps the default launch (the streaming kernel at k =
    6), the tile kernel's launch for this size (128x64, k = 5) and 128x128 at
    k = 4 all equal unblocked 128x128 tiles at k = 1."""
    N = 4096
    b = gpu.generate_fbm_terrain(N, N, seed=2)
    d = np.full((N, N), 0.5, F32)
    p = gpu.params_physical(**CFG_PARAMS)
    with gpu.Simulation(N, N, p) as s:
        s.upload(b, d)
        v0 = s.volume()
        s.step(2)
        out = s.download(flux=True)
    de, fe, _ = oracle.step(b, d, None, *p3(p), n=2)
    assert_same(out, de, fe)
    res = {}
    for kernel, k in ((2, 1), (0, 0), (3, 5), (2, 4)):
        with gpu.Simulation(N, N, p) as s:
            if kernel:
                s.set_launch(kernel, k)
            s.upload(b, d)
            s.step(200)
            res[(kernel, k)] = s.download(flux=True)
    for key in ((0, 0), (3, 5), (2, 4)):
        assert_same(res[key], res[(2, 1)]["depth"], res[(2, 1)]["flux"])
    with gpu.Simulation(N, N, p) as s:
        s.upload(b, d)
        s.step(1000)
        v1 = s.volume()
        dd = s.download(flux=False)["depth"]
        assert s.check_finite()
    assert (dd >= 0).all()
    assert abs(v1 - v0) / v0 < 1e-6, (v0, v1)


def fp64_report(label, got, r64):
    """Prints and returns (max abs, max rel with a floor of 1e-3 * max depth)
    of FP32 depth `got` against the FP64 restatement `r64`."""
    m = float(r64.max())
    err = np.abs(got.astype(np.float64) - r64)
    rel = float((err / np.maximum(np.abs(r64), 1e-3 * m)).max())
    print(f"{label}: FP32 vs FP64 max abs {err.max():.3e} m, max rel (floor 1e-3*max) "
          f"{rel:.3e}, max depth {m:.3g} m")
    return float(err.max()), rel, m


def test_full_size_4096_window_vs_restatements(gpu, oracle):
    """The 4096^2 sheet workload with the default launch: the centre 1024^2
    window after 256 steps is bit-identical to the FP32 restatement run on the
    2052^2 sub-domain that holds the window's dependency cone (2 cells per
    step); FP32 vs FP64 error reported on the same window."""
    N, window, t = 4096, 1024, 256
    b = gpu.generate_fbm_terrain(N, N, seed=2)
    d = np.full((N, N), 0.5, F32)
    p = gpu.params_physical(**CFG_PARAMS)
    w0 = (N - window) // 2
    with gpu.Simulation(N, N, p) as s:
        s.upload(b, d)
        s.step(t)
        win = s.download(flux=False)["depth"][w0:w0 + window, w0:w0 + window].copy()
    pad = 2 * t + 2
    s0, s1 = w0 - pad, w0 + window + pad
    bs = np.ascontiguousarray(b[s0:s1, s0:s1])
    ds = np.ascontiguousarray(d[s0:s1, s0:s1])
    oracle.set_threads(16)
    d32, _, _ = oracle.step(bs, ds, None, *p3(p), n=t)
    assert same_bits(win, d32[pad:pad + window, pad:pad + window])
    d64, _, _ = oracle.step(bs, ds, None, *p3(p), n=t, dtype=np.float64)
    err, _, m = fp64_report(f"4096^2 sheet, centre {window}^2 after {t} steps", win,
                            d64[pad:pad + window, pad:pad + window])
    assert err < 1e-4 * m


@pytest.mark.timeout(1200)
def test_config2_full_domain_vs_restatements(gpu, oracle, capsys):
    """The 4096^2 sheet workload (fBm seed 2 + 0.5 m, closed walls) for its
    full 1000 steps on the default launch: the whole depth and flux state is
    bit-identical to the FP32 restatement, and its FP32 error against the FP64
    restatement over the whole domain stays within the stated bounds.

    FP32 vs FP64 here is the model's own sensitivity to rounding, which the
    shaders share (they compute the same FP32 values): a thin sheet draining
    over steep terrain is dominated by wet/dry fronts, where the flux scaling
    (flowUpdate.comp:58-59) and the max(0, .) clamps switch, so a rounding
    difference moves a front by a cell.  Measured (host FP64 restatement,
    1000 steps): median abs 1.5e-6 m, p99 0.011 m, p99.9 0.055 m, max 0.80 m
    (87 of 16.8 M cells above 0.5 m; max depth 5.49 m); volume equal to
    1.6e-8 relative."""
    N, steps, chunk = 4096, 1000, 100
    b = gpu.generate_fbm_terrain(N, N, seed=2)
    d = np.full((N, N), 0.5, F32)
    p = gpu.params_physical(**CFG_PARAMS)
    with gpu.Simulation(N, N, p) as s:
        s.upload(b, d)
        s.step(steps)
        out = s.download(flux=True)
        assert s.check_finite()
    oracle.set_threads(16)
    F, K, Cc = p3(p)

    def restate(dtype):
        dd, ff = d, None
        for i in range(0, steps, chunk):
            dd, ff, _ = oracle.step(b, dd, ff, F, K, Cc, n=chunk, dtype=dtype)
            with capsys.disabled():
                print(f"  config 2 {np.dtype(dtype).name} restatement: {i + chunk}/{steps} steps",
                      flush=True)
        return dd, ff

    d32, f32 = restate(np.float32)
    assert same_bits(out["depth"], d32)
    assert same_bits(out["flux"], f32)
    del d32, f32
    d64, _ = restate(np.float64)
    err = np.abs(out["depth"].astype(np.float64) - d64)
    q50, q99, q999 = (float(np.quantile(err, q)) for q in (0.5, 0.99, 0.999))
    v0 = float(d.sum(dtype=np.float64))
    v32 = float(out["depth"].sum(dtype=np.float64))
    v64 = float(d64.sum())
    with capsys.disabled():
        print(f"  config 2, 1000 steps, FP32 vs FP64 over 4096^2: median {q50:.2e} m, p99 "
              f"{q99:.2e} m, p99.9 {q999:.2e} m, max {err.max():.3f} m (max depth "
              f"{d64.max():.2f} m); volume FP32 {abs(v32 - v0) / v0:.1e}, FP64 "
              f"{abs(v64 - v0) / v0:.1e} relative drift", flush=True)
    assert q50 < 1e-5 and q99 < 0.05 and q999 < 0.1 and err.max() < 1.5
    assert abs(v32 - v0) / v0 < 1e-6 and abs(v64 - v0) / v0 < 1e-9


def test_split_run_equals_continuous(gpu):
    b, d, f = random_state(5, 300, 200)
    p = gpu.params_physical(**CFG_PARAMS)
    full = run_gpu(gpu, b, d, f, 37, 2, 4, p, velocity=False)
    with gpu.Simulation(200, 300, p) as s:
        s.upload(b, d, f)
        s.step(20)
        mid = s.download(flux=True)
    with gpu.Simulation(200, 300, p) as s:
        s.upload(b, mid["depth"], mid["flux"])
        s.step(17)
        end = s.download(flux=True)
    assert_same(end, full["depth"], full["flux"])


def test_brush_once(gpu, oracle):
    b, d, f = random_state(8, 100, 120)
    p = gpu.params_physical(**CFG_PARAMS)
    with gpu.Simulation(120, 100, p) as s:
        s.upload(b, d, f)
        s.apply_brush(33.5, 47.25, 0.02, 32.0)
        out = s.download(flux=False)
    assert same_bits(out["depth"], oracle.brush(d, 33.5, 47.25, 0.02, 32.0))


def test_reference_layout_roundtrip(gpu, oracle):
    b, d, f = random_state(9, 64, 96)
    p = gpu.params_physical(**CFG_PARAMS)
    terrain = np.zeros((64, 96, 4), F32)
    terrain[..., 0], terrain[..., 1], terrain[..., 2], terrain[..., 3] = b, 0.3, 0.3, d
    with gpu.Simulation(96, 64, p) as s:
        s.upload_reference_layout(terrain, f)
        s.step(5)
        ref = s.download_reference_layout(gb_fill=0.3)
        out = s.download(flux=True, velocity=True)
    de, fe, ve = oracle.step(b, d, f, *p3(p), n=5, velocity=True)
    assert_same(out, de, fe, ve)
    assert same_bits(ref["terrain_rgba"][..., 0], b)
    assert same_bits(ref["terrain_rgba"][..., 3], de)
    assert (ref["terrain_rgba"][..., 1:3] == F32(0.3)).all()
    assert same_bits(ref["flow_rgba"], fe)
    # rg16f store rounds toward zero (the Mesa behaviour the fixtures pin)
    assert same_bits(ref["flowmap_rg16f"], gpu.float_to_half_rtz(ve))


def test_nonfinite_input_is_flagged(gpu):
    b, d, f = random_state(11, 64, 64)
    d[10, 20] = np.nan
    p = gpu.params_physical(**CFG_PARAMS)
    with gpu.Simulation(64, 64, p) as s:
        s.upload(b, d, f)
        s.step(3)
        assert not s.check_finite()
        assert s.check_finite()  # flag cleared by the read


def test_perform_simulation_step_clamps_like_reference(gpu):
    p = gpu.params_physical(**CFG_PARAMS)
    with gpu.Simulation(32, 32, p) as s:
        assert s.perform_simulation_step(0.05, 1.0 / 60.0) == 3
        assert s.perform_simulation_step(1.0, 1.0 / 60.0) == 10  # clamp; time dropped
        assert s.perform_simulation_step(0.0, 1.0 / 60.0) == 0
        assert s.perform_simulation_step(1e300, 1.0 / 60.0) == 10  # no integer overflow
        assert s.perform_simulation_step(0.0, 1.0 / 60.0) == 0
        for bad in (float("inf"), float("nan"), -1.0):
            with pytest.raises(gpu.TwsError):
                s.perform_simulation_step(bad, 1.0 / 60.0)


def test_reference_default_scene_steps_bitwise(gpu, oracle):
    """The drop-in default: the reference's own scene (its PRNG and noise), its
    default constants (60 steps/s, damping 0.98, acceleration 10) and its own
    out-of-image edge, 30 steps in 3 calls: bit-identical to the restatement."""
    N = 256
    b, d = gpu.generate_reference_scene(N)
    p = gpu.params_reference(boundary=gpu.BOUNDARY_ZERO)
    with gpu.Simulation(N, N, p) as s:
        s.upload(b, d)
        for n in (1, 9, 20):
            s.step(n)
        out = s.download(flux=True, velocity=True)
    de, fe, ve = oracle.step(b, d, None, *p3(p), n=30, boundary=oracle.BOUNDARY_ZERO,
                             velocity=True)
    assert_same(out, de, fe, ve)
    assert out["depth"].sum() > 0


@pytest.mark.parametrize("shape,rain", [((88, 100), (50.0, 70.0, 1e-2, 400.0)),
                                        ((168, 200), None),
                                        ((33, 47), (10.0, 30.0, 1e-2, 100.0)),
                                        ((96, 128), (60.0, 40.0, 1e-2, 300.0))])
def test_reference_dispatch_bitwise(gpu, oracle, shape, rain):
    """TWS_BOUNDARY_REFERENCE_DISPATCH: the reference's truncating dispatch
    shapes (Terrain.cpp:167,258,264) on the default launch and two others,
    then a one-shot brush: bit-identical to the restatement's DISPATCH mode.
    At 96 x 128 (multiples of 32) the tile kernels run it as the ZERO edge."""
    H, W = shape
    b, d, f = random_state(H * 7 + W, H, W)
    p = gpu.params_physical(**CFG_PARAMS, boundary=gpu.BOUNDARY_REFERENCE_DISPATCH)
    de, fe, ve = oracle.step(b, d, f, *p3(p), n=9, boundary=oracle.BOUNDARY_DISPATCH, rain=rain,
                             velocity=True)
    for kernel, k in ((0, 0), (3, 5), (1, 1)):
        out = run_gpu(gpu, b, d, f, 9, kernel, k, p, rain=rain)
        assert_same(out, de, fe, ve)
    with gpu.Simulation(W, H, p) as s:
        s.upload(b, d, f)
        s.apply_brush(W - 3.5, H - 2.25, 0.5, 64.0)
        got = s.download(flux=False)["depth"]
    bw, bh = W // 32 * 32, H // 32 * 32
    want = d.copy()
    if bw and bh:
        want[:bh, :bw] = oracle.brush(np.ascontiguousarray(d[:bh, :bw]), W - 3.5, H - 2.25, 0.5,
                                      64.0)
    assert same_bits(got, want)


@pytest.mark.parametrize("kernel,k", [(1, 1), (2, 4), (3, 5), (4, 8), (5, 6), (6, 3), (7, 8),
                                      (8, 5)])
@pytest.mark.parametrize("K", [0.0, 1e-35])
def test_closed_walls_exact_for_any_k(gpu, oracle, kernel, k, K):
    """Closed walls with K = 0 (and a K so small that the wall's head
    difference no longer outweighs a stored flux), from a state whose edge
    cells hold uploaded flux toward the walls: bit-identical to the
    restatement, which zeroes every wall-ward flux, and no water is lost."""
    b, d, f = random_state(31 + kernel, 70, 130)
    p = gpu.params_raw(0.99, K, 0.02, gpu.BOUNDARY_CLOSED)
    rain = (20.0, 60.0, 1e-2, 100.0)
    out = run_gpu(gpu, b, d, f, 7, kernel, k, p, rain=rain)
    de, fe, ve = oracle.step(b, d, f, 0.99, K, 0.02, n=7, rain=rain, velocity=True)
    assert_same(out, de, fe, ve)
    dep = float(oracle.brush(np.zeros((70, 130)), *rain, dtype=np.float64).sum()) * 7
    v0 = float(d.sum(dtype=np.float64))
    assert abs(float(out["depth"].sum(dtype=np.float64)) - (v0 + dep)) < 1e-5 * (v0 + dep)


@pytest.mark.parametrize("k", [2, 4, 5, 6, 8])
def test_stream_segments_bitwise(gpu, oracle, k):
    """The streaming kernel cuts a tall grid into many strip segments, each
    sweeping 2k rows into its neighbours; 13 steps (full launches and a tail
    of other step counts) with rain across segment seams, bitwise against the
    restatement."""
    H, W = 1500, 300
    b, d, f = random_state(400 + k, H, W)
    p = gpu.params_physical(**CFG_PARAMS)
    rain = (150.0, 700.0, 5e-3, 2500.0)
    out = run_gpu(gpu, b, d, f, 13, gpu.KERNEL_STREAM_128, k, p, rain=rain)
    de, fe, ve = oracle.step(b, d, f, *p3(p), n=13, rain=rain, velocity=True)
    assert_same(out, de, fe, ve)
