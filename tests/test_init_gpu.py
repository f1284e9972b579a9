"""Device-side initial state: the fBm terrain computed on the GPU equals the
host generator bit for bit; the water initialiser writes the sheet/column."""
import numpy as np
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
    assert np.array_equal(ref["terrain_rgba"][..., 0], host)


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
    assert np.array_equal(out0["depth"], d0) and not out0["flux"].any()
    de, fe, _ = oracle.step(b, d0, None, p.flow_friction, p.water_acceleration,
                            p.cell_area_inv_timescaled, n=17)
    assert np.array_equal(out["depth"], de) and np.array_equal(out["flux"], fe)


def test_batch_device_init(gpu):
    with gpu.Batch(4, 1 << 28) as batch:
        g = batch.add(64, 48)
        batch.init_terrain_fbm(g, 3)
        batch.init_water(g, 0.0, (10, 10, 20, 20), 5.0)
        d = batch.download(g, flux=False)["depth"]
    assert d.sum() == 5.0 * 100
