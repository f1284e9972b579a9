"""CPU restatement of the reference's default scene: its Mersenne-Twister
variant (Random.cpp:22-56), its value-noise fBm (NoiseGenerator.cpp:6-97) and
the terrain + lake of Terrain::CreateHeightmapFromNoiseAndResetSim
(Terrain.cpp:200-222).

Test infrastructure only: imported by tests/ as the checker of
tws_generate_reference_scene, never by the library.  Written in numpy float32
so that every operation rounds to float as the reference's x64 build does;
independent of the library's C++ (no shared code).

Pins (tests/test_refscene.py):
* generator: after its seeded fill the reference runs one standard MT19937
  twist, and its incremental twist + tempering produce the standard MT19937
  stream, so CPython's `random` (a C MT19937) given the same 624 words must
  yield the same 32-bit draws; the fill is checked on hand-computed words;
* noise: the k0..k7 polynomial (NoiseGenerator.cpp:83-96) against the
  trilinear form the reference keeps in a comment (NoiseGenerator.cpp:78-79),
  exact white-noise values on lattice points, the quintic fade's closed values;
* lake: the closed form at the grid centre.
Parity unpinned: the host pow(r^2, 2.0f) of Terrain.cpp:220 is taken as the
correctly rounded square (MSVC's powf is not available here).
"""
from __future__ import annotations

import numpy as np

F32 = np.float32
N_MT, M_MT = 624, 397
MATRIX_A = (0, 0x9908B0DF)
UPPER, LOWER = 0x80000000, 0x7FFFFFFF
MASK32 = 0xFFFFFFFF


def fill_table(seed: int) -> list[int]:
    """Random.cpp:25-29: the seeded fill, before the twist."""
    seed &= MASK32
    return [((seed + i * 527) if i % 2 else ((2135 + seed * 74111) * i)) & MASK32
            for i in range(N_MT)]


def init_table(seed: int) -> list[int]:
    """Random.cpp:24-38: fill, then one pass of the twist."""
    mt = fill_table(seed)
    for i in range(N_MT):
        y = (mt[i] & UPPER) | (mt[0 if i + 1 == N_MT else i + 1] & LOWER)
        mt[i] = mt[(i + M_MT) % N_MT] ^ (y >> 1) ^ MATRIX_A[y & 1]
    return mt


def draws(seed: int, n: int) -> list[int]:
    """The tempered 32-bit words of the first n Random::NextFloat calls
    (Random.cpp:42-54)."""
    mt = init_table(seed)
    idx = 0
    out = []
    for _ in range(n):
        nxt = 0 if idx + 1 == N_MT else idx + 1
        y = (mt[idx] & UPPER) | (mt[nxt] & LOWER)
        mt[idx] = mt[(idx + M_MT) % N_MT] ^ (y >> 1) ^ MATRIX_A[y & 1]
        y = mt[idx]
        idx = nxt
        y ^= y >> 11
        y ^= (y << 7) & 0x9D2C5680
        y ^= (y << 15) & 0xEFC60000
        y ^= y >> 18
        out.append(y & MASK32)
    return out


def to_float(y: int) -> np.float32:
    """Random.cpp:56: (float)(y * 4.656612874e-10 - 1.0), in double then rounded."""
    return F32(float(y) * 4.656612874e-10 - 1.0)


def white_noise(seed: int, scene_index: int = 0) -> np.ndarray:
    """NoiseGenerator.cpp:6-10: the 4096 values of the (scene_index+1)-th
    NoiseGenerator after Random::Init(seed)."""
    ys = draws(seed, 4096 * (scene_index + 1))[4096 * scene_index:]
    return np.array([to_float(y) for y in ys], F32)


def smooth(f):
    """NoiseGenerator.h:22: f^3 (f (6 f - 15) + 10), float by float."""
    f = np.asarray(f, F32)
    return f * f * f * (f * (f * F32(6.0) - F32(15.0)) + F32(10.0))


def _floor_int(a):
    """NoiseGenerator.h:24: truncation corrected for negative non-integers."""
    r = np.trunc(a).astype(np.int64)
    return r - ((a < 0) & (a - r.astype(F32) != 0)).astype(np.int64)


def noise3d(white, x, y, z, period: int, trilinear: bool = False):
    """NoiseGenerator.cpp:40-97 on float32 arrays.  trilinear=True evaluates the
    commented lerp form of NoiseGenerator.cpp:78-79 instead of the k0..k7
    polynomial (a cross-check, not bit-identical)."""
    x, y, z = (np.asarray(v, F32) for v in (x, y, z))
    period = min(int(period), 16)
    mod = period - 1
    x0, y0, z0 = _floor_int(x), _floor_int(y), _floor_int(z)
    fx, fy, fz = x - x0.astype(F32), y - y0.astype(F32), z - z0.astype(F32)
    x0, y0, z0 = (x0 % period + period) & mod, (y0 % period + period) & mod, (z0 % period + period) & mod
    x1, y1, z1 = (x0 + 1) & mod, (y0 + 1) & mod, (z0 + 1) & mod

    def s(a, b, c):
        return white[a + 16 * (b + 16 * c)]

    s000, s100, s010, s110 = s(x0, y0, z0), s(x1, y0, z0), s(x0, y1, z0), s(x1, y1, z0)
    s001, s101, s011, s111 = s(x0, y0, z1), s(x1, y0, z1), s(x0, y1, z1), s(x1, y1, z1)
    u, v, w = smooth(fx), smooth(fy), smooth(fz)
    if trilinear:
        def lrp(a, b, t):
            return a + (b - a) * t
        return lrp(lrp(lrp(s000, s100, u), lrp(s010, s110, u), v),
                   lrp(lrp(s001, s101, u), lrp(s011, s111, u), v), w)
    uv, uw, vw = u * v, u * w, v * w
    k0 = s000
    k1 = s100 - s000
    k2 = s010 - s000
    k3 = s001 - s000
    k4 = s110 - s010 - k1
    k5 = s000 - s010 - s001 + s011
    k6 = -k1 - s001 + s101
    k7 = -k4 + s001 - s101 - s011 + s111
    return k0 + k1 * u + k2 * v + k3 * w + k4 * uv + k5 * vw + k6 * uw + k7 * uv * w


def value_noise(white, x, y, z, low: int, high: int, persistence: float, periodic: bool):
    """NoiseGenerator.cpp:12-38 without the gradient."""
    x, y, z = (np.asarray(v, F32) for v in (x, y, z))
    p = F32(persistence)
    res = np.zeros(np.broadcast(x, y, z).shape, F32)
    amplitude = F32(1.0)
    frequency = F32(1 << low)
    for _ in range(low, high + 1):
        n = noise3d(white, x * frequency, y * frequency, z * frequency,
                    int(frequency) if periodic else 16)
        res = res + amplitude * (n * F32(0.5) + F32(0.5))
        amplitude = F32(amplitude * p)
        frequency = F32(frequency * F32(2.0))
    return res * F32(2.0) * (F32(1.0) - p) / (F32(1.0) - amplitude) - F32(1.0)


def scene(resolution: int, height_scale: float = 300.0, seed: int = 231656522,
          scene_index: int = 0):
    """Terrain.cpp:209-222: terrain b and lake depth d, arrays indexed [y, x]."""
    N = int(resolution)
    hs = F32(height_scale)
    white = white_noise(seed, scene_index)
    m = F32(1.0) / F32(N - 1)
    xs = np.arange(N, dtype=F32)
    X, Y = np.meshgrid(xs, xs)            # X[y, x] = x
    b = (value_noise(white, m * X, m * Y, F32(0.0), 2, 10, 0.43, True) * F32(0.5)
         + F32(0.5)) * hs
    dx, dy = X * m - F32(0.5), Y * m - F32(0.5)
    r2 = dx * dx + dy * dy
    d = np.maximum(F32(0.0), (F32(0.45) - r2 * r2 * F32(800.0)) * hs - b)
    return b.astype(F32), d.astype(F32)
