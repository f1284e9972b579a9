"""Device-side initial state: the fBm terrain computed on the GPU equals the
host generator bit for bit; the water initialiser writes the sheet/column."""
import numpy as np
from bitwise import same_bits
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape,seed,octaves,base", [((96, 160), 1, 6, 4), ((257, 129), 7, 9, 2),
                                                      ((1024, 1024), 2, 6, 4)])
def test_device_fbm_equals_host(gpu, shape, seed, octaves, base):
    H, W = shape
    host = gpu.generate_fbm_terrain(W, H, seed=seed, octaves=octaves, base_frequency=base,
                                    persistence=0.43, height_scale=100.0)
    with gpu.Simulation(W, H) as s:
        s.init_terrain_fbm(seed, octaves, base, 0.43, 100.0)
        ref = s.download_reference_layout()
    assert same_bits(ref["terrain_rgba"][..., 0], host)


def test_init_water_and_step(gpu, oracle):
    H, W = 200, 300
    p = gpu.params_physical()
    with gpu.Simulation(W, H, p) as s:
        s.init_terrain_fbm(5)
        s.init_water(0.25, (100, 50, 140, 90), 20.0)
        out0 = s.download(flux=True)
        b = s.download_reference_layout()["terrain_rgba"][..., 0].copy()
        s.step(17)
        out = s.download(flux=True)
    d0 = np.full((H, W), 0.25, np.float32)
    d0[50:90, 100:140] = 20.0
    assert same_bits(out0["depth"], d0) and not out0["flux"].any()
    de, fe, _ = oracle.step(b, d0, None, p.flow_friction, p.water_acceleration,
                            p.cell_area_inv_timescaled, n=17)
    assert same_bits(out["depth"], de) and same_bits(out["flux"], fe)


def test_batch_device_init(gpu):
    with gpu.Batch(4, 1 << 28) as batch:
        g = batch.add(64, 48)
        batch.init_terrain_fbm(g, 3)
        batch.init_water(g, 0.0, (10, 10, 20, 20), 5.0)
        d = batch.download(g, flux=False)["depth"]
    assert d.sum() == 5.0 * 100


# -- the reference's default scene on the device (tws_init_reference_scene) --

import tws_refscene as R  # noqa: E402  (oracle restatement)

R_SEED = 231656522  # Application.cpp:57


@pytest.mark.parametrize("N,scene_index,height_scale", [(16, 0, 300.0), (100, 0, 300.0),
                                                        (257, 1, 123.5), (513, 0, 300.0)])
def test_device_reference_scene_equals_restatement(gpu, N, scene_index, height_scale):
    b0, d0 = R.scene(N, height_scale, R_SEED, scene_index)
    with gpu.Simulation(N, N) as s:
        s.init_reference_scene(height_scale, R_SEED, scene_index)
        t = s.download_reference_layout()["terrain_rgba"]
        f = s.download(flux=True)["flux"]
    assert same_bits(t[..., 0], b0)
    assert same_bits(t[..., 3], d0)
    assert not f.any()


def test_device_reference_scene_equals_host_at_4096(gpu):
    N = 4096
    b0, d0 = gpu.generate_reference_scene(N)
    with gpu.Simulation(N, N) as s:
        s.init_reference_scene()
        t = s.download_reference_layout()["terrain_rgba"]
    assert same_bits(t[..., 0], b0)
    assert same_bits(t[..., 3], d0)


def test_device_reference_scene_resets_a_stepped_grid(gpu, oracle):
    """After steps that left the state in either ping-pong set, the scene
    replaces b, d and f; stepping it equals stepping the uploaded host scene."""
    N = 128
    b, d = gpu.generate_reference_scene(N)
    p = gpu.params_reference(boundary=gpu.BOUNDARY_ZERO)
    for pre in (1, 2):
        with gpu.Simulation(N, N, p) as s:
            s.set_launch(gpu.KERNEL_TILE_128x64, 4)
            s.init_terrain_fbm(9)
            s.init_water(1.0)
            for _ in range(pre):
                s.step(4)
            s.init_reference_scene()
            out0 = s.download(flux=True)
            s.step(12)
            out = s.download(flux=True, velocity=True)
        assert same_bits(out0["depth"], d) and not out0["flux"].any()
        de, fe, ve = oracle.step(b, d, None, p.flow_friction, p.water_acceleration,
                                 p.cell_area_inv_timescaled, n=12,
                                 boundary=oracle.BOUNDARY_ZERO, velocity=True)
        assert same_bits(out["depth"], de) and same_bits(out["flux"], fe)
        assert same_bits(out["velocity"], ve)


def test_device_reference_scene_rejects_bad_grids(gpu):
    with gpu.Simulation(64, 32) as s:
        with pytest.raises(gpu.TwsError):
            s.init_reference_scene()
    with gpu.Simulation(8, 8) as s:
        with pytest.raises(gpu.TwsError):
            s.init_reference_scene()
    with gpu.Simulation(32, 32) as s:
        with pytest.raises(gpu.TwsError):
            s.init_reference_scene(scene_index=-1)
        with pytest.raises(gpu.TwsError):
            s.init_reference_scene(height_scale=float("nan"))


def test_batch_device_reference_scene(gpu):
    b0, d0 = gpu.generate_reference_scene(64, 200.0, 7, 2)
    with gpu.Batch(4, 1 << 28) as batch:
        g = batch.add(64, 64)
        other = batch.add(64, 48)
        batch.init_reference_scene(g, 200.0, 7, 2)
        with pytest.raises(gpu.TwsError):
            batch.init_reference_scene(other)
        d = batch.download(g)
    assert same_bits(d["depth"], d0) and not d["flux"].any()
