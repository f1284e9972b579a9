"""The reference's default scene (tws_generate_reference_scene) against the
restatement in oracle/tws_refscene.py, and the restatement against what
pins it: CPython's MT19937, the trilinear form of the noise, closed forms."""
import random

import numpy as np
from bitwise import same_bits
import pytest

import tws_refscene as R

F32 = np.float32
SEED = 231656522  # Application.cpp:57


def test_generator_is_standard_mt19937_after_its_own_init():
    # The reference twists its seeded table once (Random.cpp:31-36), then
    # twists word by word as it draws; that is the standard MT19937 stream,
    # which CPython's C generator produces from the same 624 words.
    for seed in (SEED, 0, 1, 0xFFFFFFFF):
        r = random.Random()
        r.setstate((3, tuple(R.init_table(seed)) + (624,), None))
        assert [r.getrandbits(32) for _ in range(3000)] == R.draws(seed, 3000)
        # ... and its init twist is one standard twist of the fill
        r.setstate((3, tuple(R.fill_table(seed)) + (624,), None))
        r.getrandbits(32)
        assert list(r.getstate()[1][:624]) == R.init_table(seed)


def test_seeded_fill_words():
    # Random.cpp:27: odd words seed + 527 i, even words (2135 + 74111 seed) i, mod 2^32
    t = R.fill_table(SEED)
    assert t[0] == 0
    assert t[1] == SEED + 527
    assert t[2] == ((2135 + SEED * 74111) * 2) % 2**32
    assert t[623] == (SEED + 623 * 527) % 2**32


def test_draws_map_to_minus_one_one():
    assert R.to_float(0) == F32(-1.0)
    assert R.to_float(2**31) == F32(2**31 * 4.656612874e-10 - 1.0)
    assert abs(float(R.to_float(2**32 - 1)) - 1.0) < 1e-6
    w = R.white_noise(SEED)
    assert w.dtype == F32 and w.shape == (4096,)
    assert w.min() >= -1.0 and w.max() <= 1.0 and abs(float(w.mean())) < 0.05


def test_fade_closed_values():
    assert R.smooth(F32(0.0)) == 0.0 and R.smooth(F32(1.0)) == 1.0
    assert R.smooth(F32(0.5)) == 0.5


def test_noise_on_lattice_points_is_the_white_noise():
    w = R.white_noise(SEED)
    i, j, k = np.meshgrid(np.arange(16), np.arange(16), np.arange(16), indexing="ij")
    n = R.noise3d(w, i.astype(F32), j.astype(F32), k.astype(F32), 16)
    assert same_bits(n, w[i + 16 * (j + 16 * k)])
    # the lattice wraps with the period
    n4 = R.noise3d(w, F32(5.0), F32(6.0), F32(0.0), 4)
    assert n4 == w[1 + 16 * 2]


def test_noise_polynomial_matches_the_trilinear_form():
    # NoiseGenerator.cpp:83-96 against the lerp form kept at :78-79
    rng = np.random.default_rng(7)
    w = R.white_noise(SEED)
    p = rng.random((3, 5000)).astype(F32) * F32(40.0) - F32(8.0)
    a = R.noise3d(w, *p, 16)
    b = R.noise3d(w, *p, 16, trilinear=True)
    assert np.abs(a - b).max() < 2e-6
    assert np.abs(a).max() <= 1.0 + 1e-6


def test_single_octave_value_noise_is_the_noise():
    w = R.white_noise(SEED)
    rng = np.random.default_rng(3)
    x, y = rng.random((2, 200)).astype(F32)
    for octave in (2, 4, 6):
        f = F32(1 << octave)
        v = R.value_noise(w, x, y, F32(0.0), octave, octave, 0.43, True)
        n = R.noise3d(w, x * f, y * f, F32(0.0), int(f))
        assert np.abs(v - n).max() < 1e-6   # (n/2 + 1/2) 2 (1-p)/(1-p) - 1


def test_lake_closed_form_at_the_centre():
    N = 257  # (N - 1) / 2 = 128: the centre cell sits at r = 0
    b, d = R.scene(N)
    c = 128
    assert d[c, c] == max(F32(0.0), F32(0.45) * F32(300.0) - b[c, c])
    assert d[0, 0] == 0.0  # r^4 800 > 0.45 at the corners


@pytest.mark.parametrize("N,scene_index,height_scale", [(64, 0, 300.0), (100, 0, 300.0),
                                                        (257, 1, 123.5), (513, 0, 300.0)])
def test_library_matches_restatement(tws, N, scene_index, height_scale):
    b0, d0 = R.scene(N, height_scale, SEED, scene_index)
    b1, d1 = tws.generate_reference_scene(N, height_scale, SEED, scene_index)
    assert same_bits(b0, b1)
    assert same_bits(d0, d1)
    assert 0.0 <= b1.min() and b1.max() <= height_scale
    assert (d1 >= 0).all() and d1.max() > 0


def test_library_rejects_bad_arguments(tws):
    out = np.empty(16, F32)
    ptr = out.ctypes.data_as(tws.C.c_void_p) if hasattr(tws, "C") else out.ctypes.data
    assert tws.lib().tws_generate_reference_scene(1, 300.0, SEED, 0, ptr, None) == -1
    assert tws.lib().tws_generate_reference_scene(4, 300.0, SEED, -1, ptr, None) == -1
    assert tws.lib().tws_generate_reference_scene(4, 300.0, SEED, 0, None, None) == -1
