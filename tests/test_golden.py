"""The CPU restatement against the reference's own shaders.

tests/golden/*.npz hold inputs and outputs of the unmodified reference GLSL
shaders (flowUpdate.comp, flowApply.comp, waterBrush.comp) executed on Mesa
llvmpipe by tests/golden/make_goldens.py; manifest.json records the shader
hashes, renderer and constants.  Bar: bit-identical depth, flux and FlowMap
(rg16f bits) at every stored checkpoint.
"""
import json
from pathlib import Path

import numpy as np
from bitwise import same_bits
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden"
MANIFEST = json.loads((GOLDEN / "manifest.json").read_text())
CASES = sorted(MANIFEST["cases"])
F32 = np.float32


def half_rtz_numpy(x):
    """float32 -> float16 bits rounded toward zero, via numpy's nearest-even
    conversion stepped back by one ulp where it rounded away from zero."""
    x = np.asarray(x, F32)
    with np.errstate(over="ignore"):
        h = x.astype(np.float16)
    hb = h.view(np.uint16).astype(np.int64)
    away = (np.abs(h.astype(np.float64)) > np.abs(x.astype(np.float64))) & (h != 0)
    return np.where(away, hb - 1, hb).astype(np.uint16)


def load(name):
    z = np.load(GOLDEN / f"{name}.npz")
    return {k: z[k] for k in z.files}


def test_manifest_names_the_reference_shaders():
    assert {"flowUpdate.comp", "flowApply.comp", "waterBrush.comp"} <= set(MANIFEST["shaders"])
    assert "llvmpipe" in MANIFEST["renderer"]
    assert len(CASES) >= 5


@pytest.mark.parametrize("name", CASES)
def test_oracle_matches_reference_shaders(oracle, name):
    case = MANIFEST["cases"][name]
    z = load(name)
    F, K, Cc = case["F_value"], case["K_value"], case["C_value"]
    assert f"0x{np.float32(F).view(np.uint32):08x}" == case["F"]
    boundary = {"closed": oracle.BOUNDARY_CLOSED, "zero": oracle.BOUNDARY_ZERO,
                "dispatch": oracle.BOUNDARY_DISPATCH}[case["boundary"]]
    for n in case["steps"]:
        d, f, v = oracle.step(z["b"], z["d0"], z.get("f0"), F, K, Cc, n=n, boundary=boundary,
                              rain=case["rain"], velocity=True)
        assert same_bits(d, z[f"d_{n}"]), f"{name}: depth differs after {n} steps"
        assert same_bits(f, z[f"f_{n}"]), f"{name}: flux differs after {n} steps"
        assert same_bits(half_rtz_numpy(v), z[f"v16_{n}"]), f"{name}: FlowMap differs"


def test_half_rtz_library_matches_numpy(tws):
    rng = np.random.default_rng(0)
    x = np.concatenate([
        (rng.standard_normal(200000) * 10.0 ** rng.uniform(-9, 6, 200000)).astype(F32),
        np.array([0.0, -0.0, 1.0, -1.0, 65504.0, 65519.0, 65520.0, 70000.0, -70000.0,
                  6.1e-5, 5.96e-8, 2.9e-8, 1e-40, np.inf, -np.inf], F32),
    ])
    assert same_bits(tws.float_to_half_rtz(x), half_rtz_numpy(x))
    nan = tws.float_to_half_rtz(np.array([np.nan], F32))
    assert np.isnan(nan.view(np.float16)[0])


def test_fixtures_are_not_trivial():
    z = load("config1_256_rain")
    assert z["d_1000"].sum(dtype=np.float64) > 50.0       # ~1000 * 0.8 m^3 of rain
    assert np.count_nonzero(z["f_1000"]) > 1000
    z = load("random64x96_zero_edge")
    assert z["d_20"].sum(dtype=np.float64) < z["d0"].sum(dtype=np.float64)  # drained
    # the reference's dispatch shapes leave the tail of a 100 x 88 grid as it was
    z = load("random100x88_dispatch")
    assert same_bits(z["d_25"][80:], z["d0"][80:])
    assert same_bits(z["d_25"][:, 96:], z["d0"][:, 96:])
    assert same_bits(z["f_25"][:, 96:], z["f0"][:, 96:])
    assert not z["v16_25"][:, 96:].any()
    assert not same_bits(z["d_25"][64:80, :96], z["d0"][64:80, :96])


WINDOWS = sorted(MANIFEST.get("window_cases", {}))


def test_window_case_present():
    assert "config2_4096_window64" in WINDOWS


@pytest.mark.parametrize("name", WINDOWS)
def test_oracle_matches_reference_window(tws, oracle, name):
    """The headline configuration (4096^2 fBm + 0.5 m sheet, closed walls)
    after 64 steps of the reference shaders, centre window: the restatement
    run on a sub-domain that holds the window's whole dependency cone (2 cells
    per step) plus margin reproduces it bit for bit."""
    case = MANIFEST["window_cases"][name]
    z = load(name)
    N = case["shape"][0]
    (n,) = case["steps"]
    y0, x0, wh, ww = case["window"]
    b = tws.generate_fbm_terrain(N, N, seed=2, octaves=6, base_frequency=4, persistence=0.43,
                                 height_scale=100.0)
    import hashlib
    assert hashlib.sha256(b.tobytes()).hexdigest() == case["terrain_sha256"]
    m = 2 * n + 64
    sub = (slice(y0 - m, y0 + wh + m), slice(x0 - m, x0 + ww + m))
    bs = np.ascontiguousarray(b[sub])
    d, f, v = oracle.step(bs, np.full(bs.shape, 0.5, F32), None, case["F_value"],
                          case["K_value"], case["C_value"], n=n, velocity=True)
    w = (slice(m, m + wh), slice(m, m + ww))
    assert same_bits(bs[w], z["b_window"])
    assert same_bits(d[w], z[f"d_{n}"])
    assert same_bits(f[w], z[f"f_{n}"])
    assert same_bits(half_rtz_numpy(v[w]), z[f"v16_{n}"])
    assert z[f"f_{n}"].any() and not same_bits(z[f"d_{n}"], np.full((wh, ww), 0.5, F32))
