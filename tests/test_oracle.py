"""Pins of the CPU restatement (oracle/tws_oracle.c) that need no GPU.

Each check is chosen so that a plausible transcription mistake (a dropped
term, a wrong sign, a swapped channel or neighbour index, a missing clamp)
fails it: closed-form first steps with exactly representable constants,
invariants (lake at rest, non-negativity, volume conservation, x-mirror
equivariance), an independent whole-array numpy restatement compared bit for
bit, and FP32-vs-FP64 agreement.  The llvmpipe golden fixtures of the
unmodified shaders are checked in test_golden.py.
"""
import os
import shutil
import subprocess
from fractions import Fraction
from pathlib import Path

import numpy as np
from bitwise import same_bits
import pytest

ROOT = Path(__file__).resolve().parent.parent

F32 = np.float32


def test_two_cells_closed_form(oracle):
    # 1 x 2 grid, water only in the left cell.  K = 1/4, C = 1/2, F = 1/2 are
    # exact, so every intermediate is exact: f' = h*K = 1/4 toward +x, the
    # scaling sum 1/4 * 1/2 = 1/8 < d = 1 leaves it alone, and the depth update
    # moves 1/4 * 1/2 = 1/8 across the pipe.
    b = np.zeros((1, 2), F32)
    d = np.array([[1.0, 0.0]], F32)
    d1, f1, v1 = oracle.step(b, d, F=0.5, K=0.25, Cc=0.5, n=1, velocity=True)
    assert d1.tolist() == [[0.875, 0.125]]
    assert f1[0, 0].tolist() == [0.25, 0.0, 0.0, 0.0]
    assert f1[0, 1].tolist() == [0.0, 0.0, 0.0, 0.0]
    # v.x = (X1.y - f.+x) - (X0.x - f.-x): -1/4 in the left cell, -1/4 in the
    # right cell (it points against the net transport, flowApply.comp:45).
    assert v1[0, 0].tolist() == [-0.25, 0.0]
    assert v1[0, 1].tolist() == [-0.25, 0.0]


def test_second_step_uses_friction(oracle):
    # Step 2 of the 1 x 2 case: f' = f*F + g*K with f = 1/4, F = 1/2,
    # g = 0.875 - 0.125 = 0.75 -> 1/8 + 3/16 = 5/16; d moves by 5/32.
    b = np.zeros((1, 2), F32)
    d = np.array([[1.0, 0.0]], F32)
    d2, f2, _ = oracle.step(b, d, F=0.5, K=0.25, Cc=0.5, n=2)
    assert f2[0, 0, 0] == F32(5 / 16)
    assert d2.tolist() == [[0.875 - 5 / 32, 0.125 + 5 / 32]]


def test_three_by_three_spreads_evenly(oracle):
    b = np.zeros((3, 3), F32)
    d = np.zeros((3, 3), F32)
    d[1, 1] = 1.0
    d1, f1, v1 = oracle.step(b, d, F=0.5, K=0.25, Cc=0.5, n=1, velocity=True)
    expect = np.array([[0, 0.125, 0], [0.125, 0.5, 0.125], [0, 0.125, 0]], F32)
    assert same_bits(d1, expect)
    assert f1[1, 1].tolist() == [0.25, 0.25, 0.25, 0.25]
    assert np.count_nonzero(f1) == 4
    # directed flow: zero at the centre, pointing back toward the centre
    # around it (+x neighbour: v.x = (0 - 0) - (1/4 - 0) = -1/4)
    assert v1[1, 1].tolist() == [0.0, 0.0]
    assert v1[1, 2].tolist() == [-0.25, 0.0]
    assert v1[1, 0].tolist() == [0.25, 0.0]
    assert v1[2, 1].tolist() == [0.0, -0.25]
    assert v1[0, 1].tolist() == [0.0, 0.25]


def test_scaling_keeps_depth_non_negative(oracle):
    # A tall column with little water: the unscaled outflow (4 * 10.25 * K *
    # C = 5.125) exceeds d = 0.25, so flowUpdate.comp:58-59 scales it by
    # d / total and the cell drains to (almost exactly) zero.
    b = np.zeros((3, 3), F32)
    b[1, 1] = 10.0
    d = np.zeros((3, 3), F32)
    d[1, 1] = 0.25
    d1, f1, _ = oracle.step(b, d, F=0.5, K=0.25, Cc=0.5, n=1)
    s = F32(0.25) / F32(F32(4 * 2.5625) * F32(0.5))
    assert f1[1, 1, 0] == F32(F32(2.5625) * s)
    assert 0.0 <= d1[1, 1] <= 1e-7
    assert abs(float(np.sum(d1, dtype=np.float64)) - 0.25) < 1e-7
    # without the clamp the centre would go negative by 4.875
    dd, _, _ = oracle.step(b, d, F=0.5, K=0.25, Cc=0.5, n=1, dtype=np.float64)
    assert dd[1, 1] >= 0.0


def test_lake_at_rest_is_exact(oracle):
    rng = np.random.default_rng(7)
    b = (rng.integers(0, 32, size=(23, 31)) / 8.0).astype(F32)
    d = (F32(8.0) - b).astype(F32)
    f = np.zeros(b.shape + (4,), F32)
    for boundary in (oracle.BOUNDARY_CLOSED,):
        d1, f1, _ = oracle.step(b, d, f, F=0.9999, K=0.1962, Cc=0.02, n=50, boundary=boundary)
        assert same_bits(d1, d)
        assert not np.any(f1)


def _random_case(seed, H=37, W=29, wet=0.7):
    rng = np.random.default_rng(seed)
    b = (rng.random((H, W)) * 5.0).astype(F32)
    d = (rng.random((H, W)) * 2.0 * (rng.random((H, W)) < wet)).astype(F32)
    f = (rng.random((H, W, 4)) * 0.5).astype(F32)
    return b, d, f


def test_non_negative_and_volume_conserved(oracle):
    b, d, f = _random_case(1)
    v0 = float(np.sum(d, dtype=np.float64))
    d64, f64, _ = oracle.step(b, d, f, n=400, dtype=np.float64)
    d32, f32, _ = oracle.step(b, d, f, n=400)
    assert (d32 >= 0).all() and (f32 >= 0).all() and (d64 >= 0).all()
    assert abs(float(d64.sum()) - v0) / v0 < 1e-12
    assert abs(float(np.sum(d32, dtype=np.float64)) - v0) / v0 < 1e-5


def test_zero_boundary_drains(oracle):
    # Reference-edge mode: outside reads as zero height, so water leaves.
    b = np.zeros((8, 8), F32)
    d = np.ones((8, 8), F32)
    d1, _, _ = oracle.step(b, d, n=20, boundary=oracle.BOUNDARY_ZERO)
    assert d1.sum() < 64.0
    dc, _, _ = oracle.step(b, d, n=20, boundary=oracle.BOUNDARY_CLOSED)
    assert same_bits(dc, d)  # flat and closed: nothing moves


def test_x_mirror_is_bitwise(oracle):
    b, d, f = _random_case(2)
    rain = (11.0, 17.0, 1e-3, 64.0)
    d1, f1, v1 = oracle.step(b, d, f, n=60, rain=rain, velocity=True)
    W = b.shape[1]
    fm = f[:, ::-1][..., [1, 0, 2, 3]]
    rain_m = (float(W - 1) - 11.0, 17.0, 1e-3, 64.0)
    d2, f2, v2 = oracle.step(b[:, ::-1], d[:, ::-1], fm, n=60, rain=rain_m, velocity=True)
    assert same_bits(d2[:, ::-1], d1)
    assert same_bits(f2[:, ::-1][..., [1, 0, 2, 3]], f1)
    # negation turns a +0 velocity into -0: equal values, not equal bits
    assert np.array_equal(-v2[:, ::-1, 0], v1[..., 0])
    assert same_bits(v2[:, ::-1, 1], v1[..., 1])


def _numpy_step(b, d, f, F, K, Cc, n, closed=True):
    """Independent whole-array restatement (shifted-array form) in float32."""
    F, K, Cc = F32(F), F32(K), F32(Cc)
    H, W = b.shape
    d = d.copy()
    f = f.copy()
    for _ in range(n):
        h = (d + b).astype(F32)
        hp = np.full((H + 2, W + 2), np.nan, F32)
        hp[1:-1, 1:-1] = h
        nb = [hp[1:-1, 2:], hp[1:-1, :-2], hp[2:, 1:-1], hp[:-2, 1:-1]]
        o = []
        for c in range(4):
            t = (f[..., c] * F).astype(F32) + ((h - nb[c]).astype(F32) * K).astype(F32)
            t = np.where(np.isnan(t), F32(0), np.maximum(F32(0), t)).astype(F32)
            o.append(t)
        tot = (((o[0] + o[1]).astype(F32) + (o[2] + o[3]).astype(F32)).astype(F32) * Cc).astype(F32)
        sc = tot > d
        with np.errstate(divide="ignore", invalid="ignore"):
            s = np.where(sc, d / np.where(sc, tot, F32(1)), F32(1)).astype(F32)
        o = [np.where(sc, (oc * s).astype(F32), oc) for oc in o]
        op = [np.zeros((H + 2, W + 2), F32) for _ in range(4)]
        for c in range(4):
            op[c][1:-1, 1:-1] = o[c]
        x1y = op[1][1:-1, 2:]
        x0x = op[0][1:-1, :-2]
        y1w = op[3][2:, 1:-1]
        y0z = op[2][:-2, 1:-1]
        inflow = (((x1y + x0x).astype(F32) + y1w).astype(F32) + y0z).astype(F32)
        out = ((((o[0] + o[1]).astype(F32) + o[2]).astype(F32) + o[3])).astype(F32)
        d = np.maximum(F32(0), (d + ((inflow - out).astype(F32) * Cc).astype(F32)).astype(F32))
        f = np.stack(o, axis=-1).astype(F32)
    return d, f


@pytest.mark.parametrize("seed", [3, 4, 5])
def test_matches_independent_numpy_restatement(oracle, seed):
    b, d, f = _random_case(seed)
    de, fe = _numpy_step(b, d, f, 0.99989975, 0.1962, 0.02, 40)
    do, fo, _ = oracle.step(b, d, f, F=0.99989975, K=0.1962, Cc=0.02, n=40)
    assert same_bits(do, de)
    assert same_bits(fo, fe)


def test_fp32_tracks_fp64(oracle):
    b, d, f = _random_case(6, 64, 64)
    # relative error with a floor of 1e-3 * max depth (|dd|/d is meaningless
    # in nearly dry cells); measured: 9e-7 after 1 step, 1.1e-4 after 200.
    d1_32, _, _ = oracle.step(b, d, f, n=1)
    d1_64, _, _ = oracle.step(b, d, f, n=1, dtype=np.float64)
    m = float(d1_64.max())
    assert (np.abs(d1_32 - d1_64) / np.maximum(np.abs(d1_64), 1e-3 * m)).max() < 2e-6
    d32, _, _ = oracle.step(b, d, f, n=200)
    d64, _, _ = oracle.step(b, d, f, n=200, dtype=np.float64)
    m = float(d64.max())
    rel = np.abs(d32 - d64) / np.maximum(np.abs(d64), 1e-3 * m)
    assert rel.max() < 5e-4
    assert np.abs(d32 - d64).max() / m < 2e-5


def test_brush_closed_form(oracle):
    # waterBrush.comp:26-28 with an integer centre and R^2: each texel gets
    # max(0, 1 - r^2/R^2) * I exactly in rational arithmetic.
    H, W, cx, cy, I, R2 = 40, 50, 20.0, 17.0, 0.5, 64.0
    d = oracle.brush(np.zeros((H, W), np.float64), cx, cy, I, R2, dtype=np.float64)
    total = Fraction(0)
    for y in range(H):
        for x in range(W):
            r2 = (int(cx) - x) ** 2 + (int(cy) - y) ** 2
            w = max(Fraction(0), 1 - Fraction(r2) / Fraction(int(R2)))
            total += w * Fraction(I)
            assert abs(d[y, x] - float(w * Fraction(I))) <= 1e-15
    assert abs(float(d.sum()) - float(total)) < 1e-12
    d32 = oracle.brush(np.zeros((H, W), F32), cx, cy, I, R2)
    assert d32[17, 20] == F32(0.5)
    assert d32[17, 28] == 0.0 and d32[17, 27] > 0.0
    assert np.allclose(d32, d, rtol=0, atol=1e-7)


def test_rain_volume_matches_deposits(oracle):
    # closed walls: volume after n steps = n * (brush total), in FP64
    H, W = 48, 48
    rng = np.random.default_rng(9)
    b = (rng.random((H, W)) * 3.0).astype(F32)
    d = np.zeros((H, W), F32)
    rain = (24.0, 24.0, 1e-3, 256.0)
    dep = float(oracle.brush(np.zeros((H, W), np.float64), *rain, dtype=np.float64).sum())
    d64, _, _ = oracle.step(b, d, n=100, rain=rain, dtype=np.float64)
    assert abs(float(d64.sum()) - 100 * dep) / (100 * dep) < 1e-12
    d32, _, _ = oracle.step(b, d, n=100, rain=rain)
    assert abs(float(np.sum(d32, dtype=np.float64)) - 100 * dep) / (100 * dep) < 1e-6


def test_params_reference_defaults(oracle):
    # Terrain.cpp:22-30,95 defaults: 60 steps/s, damping 0.98, accel 10, l = 1.
    F, K, Cc = oracle.params_reference(60.0, 0.98, 10.0, 1024.0, 1024)
    dt = float(F32(1.0) / F32(60.0))
    assert K == float(F32(dt * 10.0 * 1.0))
    assert Cc == float(F32(dt / 1.0))
    assert abs(F - 0.98 ** dt) <= 1.2e-7  # powf: within an ulp of the exact power
    F2, K2, C2 = oracle.params_physical(0.02, 9.81, 1.0, 1.0, 0.995)
    assert K2 == float(F32(0.02 * 1.0 * float(F32(9.81)) / 1.0)) or abs(K2 - 0.1962) < 1e-7
    assert C2 == float(F32(float(F32(0.02))))
    assert abs(F2 - 0.995 ** 0.02) < 1e-7


def test_dispatch_mode_freezes_the_uncovered_tail(oracle):
    """BOUNDARY_DISPATCH: the reference's truncating dispatch shapes.  At
    multiples of 32 it is the ZERO edge; otherwise columns >= 16*(W/16) and
    rows >= 16*(H/16) keep depth and flux, and the rain reaches only
    [0, 32*(W/32)) x [0, 32*(H/32))."""
    rng = np.random.default_rng(5)
    F, K, Cc = oracle.params_physical()
    for H, W in ((64, 96), (70, 90), (40, 33)):
        b = (rng.random((H, W)) * 4).astype(np.float32)
        d = (rng.random((H, W)) * 2).astype(np.float32)
        f = (rng.random((H, W, 4)) * 0.3).astype(np.float32)
        rain = (W / 2.0, H - 3.0, 1e-2, 900.0)
        dz, fz, _ = oracle.step(b, d, f, F, K, Cc, n=7, boundary=oracle.BOUNDARY_ZERO, rain=rain)
        dd, fd, vd = oracle.step(b, d, f, F, K, Cc, n=7, boundary=oracle.BOUNDARY_DISPATCH,
                                 rain=rain, velocity=True)
        if H % 32 == 0 and W % 32 == 0:
            assert same_bits(dz, dd) and same_bits(fz, fd)
            continue
        uh, uw = H // 16 * 16, W // 16 * 16
        assert same_bits(dd[uh:], d[uh:]) and same_bits(dd[:, uw:], d[:, uw:])
        assert same_bits(fd[uh:], f[uh:]) and same_bits(fd[:, uw:], f[:, uw:])
        assert not vd[uh:].any() and not vd[:, uw:].any()
        assert not same_bits(dd[:uh, :uw], d[:uh, :uw])
        # the covered cells read the frozen tail as ordinary neighbours: a lake
        # at rest stays at rest next to it, while the image edge at x = -1,
        # y = -1 (read as 0) drains the lake there, 2 cells per step
        flat = np.zeros((H, W), np.float32)
        d1, f1, _ = oracle.step(flat, np.ones((H, W), np.float32), None, F, K, Cc, n=3,
                                boundary=oracle.BOUNDARY_DISPATCH)
        inner = (slice(8, uh), slice(8, uw))
        assert not f1[inner].any() and (d1[inner] == 1.0).all()
        assert f1[:2, :uw].any() and (d1[0, :uw] < 1.0).all()


SANITIZED_DRIVER = r"""
#include <stdio.h>
#include <stdlib.h>
int oracle_step_f32(int W, int H, const float* b, float* d, float* f, float F, float K, float C,
                    int boundary, float ghost_alpha, const float* rain, int nsteps, float* v);
int oracle_step_f64(int W, int H, const double* b, double* d, double* f, double F, double K,
                    double C, int boundary, double ghost_alpha, const float* rain, int nsteps,
                    double* v);
int oracle_mip_levels(int W, int H);
void oracle_mips(int W, int H, const float* rgba0, float* out);
void oracle_set_threads(int n);
static float frand(unsigned* s) { *s = *s * 1664525u + 1013904223u; return (*s >> 8) / 16777216.0f; }
int main(void) {
    const int shapes[][2] = {{1, 1}, {1, 7}, {5, 1}, {2, 3}, {17, 33}, {40, 33}, {70, 90}};
    const float rain[4] = {3.5f, 2.25f, 1e-2f, 64.0f};
    unsigned seed = 1;
    oracle_set_threads(2);
    for (unsigned s = 0; s < sizeof shapes / sizeof shapes[0]; ++s) {
        const int H = shapes[s][0], W = shapes[s][1];
        const size_t n = (size_t)W * H;
        float *b = malloc(n * 4), *d = malloc(n * 4), *f = malloc(n * 16), *v = malloc(n * 8);
        double *bd = malloc(n * 8), *dd = malloc(n * 8), *fd = malloc(n * 32), *vd = malloc(n * 16);
        for (int boundary = 0; boundary < 3; ++boundary) {
            for (size_t i = 0; i < n; ++i) {
                b[i] = 4 * frand(&seed); d[i] = 2 * frand(&seed); bd[i] = b[i]; dd[i] = d[i];
                for (int c = 0; c < 4; ++c) { f[4 * i + c] = 0.3f * frand(&seed); fd[4 * i + c] = f[4 * i + c]; }
            }
            if (oracle_step_f32(W, H, b, d, f, 0.99f, 0.1962f, 0.02f, boundary, 0.0f, rain, 9, v) ||
                oracle_step_f64(W, H, bd, dd, fd, 0.99, 0.1962, 0.02, boundary, 0.0, rain, 9, vd)) {
                printf("step failed\n"); return 1;
            }
        }
        const int levels = oracle_mip_levels(W, H);
        size_t total = 0;
        for (int l = 1, w = W, h = H; l < levels; ++l) {
            w = w > 1 ? w / 2 : 1; h = h > 1 ? h / 2 : 1; total += (size_t)w * h * 4;
        }
        float *rgba = malloc(n * 16), *mips = malloc((total ? total : 1) * 4);
        for (size_t i = 0; i < n * 4; ++i) rgba[i] = frand(&seed);
        oracle_mips(W, H, rgba, mips);
        free(rgba); free(mips);
        free(b); free(d); free(f); free(v); free(bd); free(dd); free(fd); free(vd);
    }
    printf("OK\n");
    return 0;
}
"""


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_oracle_under_sanitizers(tmp_path):
    """The C restatement built with AddressSanitizer and UndefinedBehaviorSanitizer
    (host code only): FP32 and FP64 steps with rain on degenerate and ragged
    shapes in every edge mode, and the mip chain, without a report."""
    src = tmp_path / "driver.c"
    src.write_text(SANITIZED_DRIVER)
    exe = tmp_path / "driver"
    odir = ROOT / "oracle"
    subprocess.run(["gcc", "-O1", "-g", "-std=c11", "-ffp-contract=off", "-fopenmp",
                    "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
                    f"-I{odir}", str(odir / "tws_oracle.c"), str(src), "-o", str(exe), "-lm"],
                   check=True)
    env = dict(os.environ, OMP_NUM_THREADS="2", ASAN_OPTIONS="detect_leaks=1")
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300, env=env)
    assert out.returncode == 0 and out.stdout.strip() == "OK", out.stdout + out.stderr[-3000:]
