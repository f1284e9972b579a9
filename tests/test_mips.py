"""Terrain mip chain: the restatement against llvmpipe's glGenerateMipmap
fixtures and against closed forms (CPU only).

The reference regenerates the TerrainData mip chain after each frame with
steps (Terrain.cpp:270-276 -> Texture2D::GenMipMaps, Texture2D.cpp:64-68);
the chain length follows Texture.cpp:28-44.  Fixtures: tests/golden/mips_*.npz,
written by tests/golden/make_goldens.py from Mesa llvmpipe.
"""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pytest

GOLD = Path(__file__).resolve().parent / "golden"
F32 = np.float32
POT = ["mips_fbm128", "mips_rect64x16", "mips_column1x8"]
NPOT = ["mips_npot100x60"]
# llvmpipe takes non-halving sample positions from its rasteriser's
# interpolation (rounded by up to ~1e-4 texel at 1000 texels); on the 100x60
# fixture that moves a level by at most 6.5e-7 of the level-0 value range.
NPOT_TOL = 1e-5


def rgba(b, d, gb=0.3):
    t = np.zeros(b.shape + (4,), F32)
    t[..., 0], t[..., 1], t[..., 2], t[..., 3] = b, gb, gb, d
    return t


def fixture(name):
    z = np.load(GOLD / f"{name}.npz")
    n = len([k for k in z.files if k.startswith("level")])
    return z["b"], z["d"], [z[f"level{l}"] for l in range(1, n + 1)]


def test_manifest_lists_mip_fixtures():
    m = json.loads((GOLD / "manifest.json").read_text())
    assert set(POT + NPOT) <= set(m["mip_cases"])
    for name in POT + NPOT:
        b, d, levels = fixture(name)
        assert m["mip_cases"][name]["levels"] == len(levels) + 1
        assert m["mip_cases"][name]["shape"] == list(b.shape)


@pytest.mark.parametrize("W,H,n", [(1024, 1024, 11), (1000, 1000, 10), (1, 1, 1), (3, 1, 2),
                                   (64, 16, 7), (16384, 16384, 15), (2, 4096, 13)])
def test_level_count_matches_reference_rule(oracle, W, H, n):
    # Texture::ConvertMipMapSettingToActualCount halves every dimension until
    # all reach 0 (Texture.cpp:28-44): floor(log2(max(W, H))) + 1
    assert oracle.mip_levels(W, H) == n == int(np.floor(np.log2(max(W, H)))) + 1


@pytest.mark.parametrize("name", POT)
def test_oracle_matches_llvmpipe_bitwise(oracle, name):
    b, d, gold = fixture(name)
    ours = oracle.mips(rgba(b, d))
    assert len(ours) == len(gold)
    for l, (a, g) in enumerate(zip(ours, gold), 1):
        assert a.shape == g.shape, l
        assert np.array_equal(a.view(np.uint32), g.view(np.uint32)), f"level {l}"


@pytest.mark.parametrize("name", NPOT)
def test_oracle_matches_llvmpipe_npot_within_position_rounding(oracle, name):
    b, d, gold = fixture(name)
    ours = oracle.mips(rgba(b, d))
    span = max(np.ptp(b), np.ptp(d))
    for l, (a, g) in enumerate(zip(ours, gold), 1):
        assert a.shape == g.shape
        assert np.max(np.abs(a - g)) <= NPOT_TOL * span, f"level {l}"
        assert np.all(a[..., 1:3] == F32(0.3))


def test_constant_texture_stays_constant(oracle):
    t = np.full((24, 40, 4), F32(7.25))
    for a in oracle.mips(t):
        assert np.all(a == F32(7.25))


def test_pot_level_is_block_mean_when_exact(oracle):
    # integers: every partial sum and half is exact, so the bilinear blit at
    # t = 1/2 equals the 2x2 box mean; checks the index mapping per level
    rng = np.random.default_rng(3)
    t = (rng.integers(-64, 64, (32, 16, 4)) * 4).astype(F32)
    prev = t.astype(np.float64)
    for a in oracle.mips(t):
        h, w = a.shape[:2]
        if prev.shape[0] == 1:
            box = 0.5 * (prev[:, 0::2] + prev[:, 1::2])
        elif prev.shape[1] == 1:
            box = 0.5 * (prev[0::2] + prev[1::2])
        else:
            box = 0.25 * (prev[0::2, 0::2] + prev[0::2, 1::2] + prev[1::2, 0::2] + prev[1::2, 1::2])
        assert box.shape[:2] == (h, w)
        np.testing.assert_allclose(a, box, rtol=0, atol=2.0 ** -20 * np.max(np.abs(box)) + 1e-30)
        prev = a.astype(np.float64)


def test_npot_sample_position_on_linear_ramp(oracle):
    # bilinear interpolation reproduces a linear function: texel i of a level
    # of width wd over a source of width ws holds the ramp at
    # u = (i + 0.5) * ws / wd - 0.5 (clamped to [0, ws - 1])
    W, H = 125, 7
    x = np.arange(W, dtype=np.float64)
    t = np.zeros((H, W, 4), F32)
    t[..., 0] = (3.0 * x + 1.0).astype(F32)
    a = oracle.mips(t)[0]
    wd = W // 2
    u = np.clip((np.arange(wd) + 0.5) * W / wd - 0.5, 0, W - 1)
    np.testing.assert_allclose(a[0, :, 0], 3.0 * u + 1.0, rtol=2e-6)
    # a planted mistake in the position (the halving formula u = 2i + 1/2) would fail
    assert not np.allclose(a[0, :, 0], 3.0 * (2 * np.arange(wd) + 0.5) + 1.0, rtol=1e-4)


def test_chain_size_matches_oracle(tws):
    for W, H in [(256, 256), (100, 60), (1, 8), (4096, 4096), (3, 1)]:
        levels, floats = tws.mip_chain_size(W, H)
        want = sum(4 * max(1, W >> l) * max(1, H >> l) for l in range(1, levels))
        assert levels == int(np.floor(np.log2(max(W, H)))) + 1
        assert floats == want
    with pytest.raises(tws.TwsError):
        tws.mip_chain_size(0, 5)
