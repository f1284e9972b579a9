"""The C++ host mirror of the reference's Terrain class (include/tws_terrain.hpp),
driven by tests/cpp/test_terrain.cpp (built by __graft_entry__.build())."""
import subprocess
from pathlib import Path

import numpy as np
from bitwise import same_bits
import pytest

ROOT = Path(__file__).resolve().parent.parent
EXE = ROOT / "tests" / "cpp" / "_build" / "test_terrain"


def build_cpp_tests():
    """Compiles the C++ test drivers against include/ and libtws.so."""
    out = EXE.parent
    out.mkdir(exist_ok=True)
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", f"-I{ROOT / 'include'}",
                    str(ROOT / "tests" / "cpp" / "test_terrain.cpp"), "-o", str(EXE),
                    f"-L{ROOT / 'terrainwatersim_amd'}", "-ltws",
                    "-Wl,-rpath,$ORIGIN/../../../terrainwatersim_amd"], check=True)
    return EXE


def test_cpp_mirror_compiles():
    assert build_cpp_tests().exists()


@pytest.mark.gpu
@pytest.mark.parametrize("N", [256, 200])
def test_cpp_mirror_matches_c_abi_and_oracle(gpu, oracle, tmp_path, N):
    """The mirror's frame sequence (3 steps, a brush, 10 steps, a damping
    change, 1 step) equals the same calls through the C ABI, and the CPU
    restatement with the reference's own edge: ZERO at 256 (a multiple of
    32), the truncating dispatch shapes at 200 (Terrain.cpp:167,258,264)."""
    if not EXE.exists():
        pytest.fail("tests/cpp/_build/test_terrain missing: run __graft_entry__.build()")
    pre = str(tmp_path / "t_")
    res = subprocess.run([str(EXE), str(N), pre], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0 and res.stdout.strip().endswith("OK"), res.stdout + res.stderr
    load = lambda n, *shape: np.fromfile(pre + n + ".f32", np.float32).reshape(shape)
    b0, d0 = load("b0", N, N), load("d0", N, N)
    d_gpu, f_gpu = load("d", N, N), load("f", N, N, 4)
    F1, K, Cc = oracle.params_reference(60.0, 0.98, 10.0, 1024.0, N)
    F2, _, _ = oracle.params_reference(60.0, 0.9, 10.0, 1024.0, N)
    bnd = oracle.BOUNDARY_ZERO if N % 32 == 0 else oracle.BOUNDARY_DISPATCH
    d, f, _ = oracle.step(b0, d0, None, F1, K, Cc, n=3, boundary=bnd)
    # ApplyRadialWaterBrush(1124, -20, 2): the world position wraps into the grid
    u = np.float32(1124.0) / np.float32(1024.0)
    v = np.float32(-20.0) / np.float32(1024.0)
    cx = (u - np.floor(u)) * np.float32(N)
    cy = (v - np.floor(v)) * np.float32(N)
    bw = N if bnd == oracle.BOUNDARY_ZERO else N // 32 * 32
    d[:bw, :bw] = oracle.brush(np.ascontiguousarray(d[:bw, :bw]), float(cx), float(cy), 2.0, 32.0)
    d, f, _ = oracle.step(b0, d, f, F1, K, Cc, n=10, boundary=bnd)
    d, f, _ = oracle.step(b0, d, f, F2, K, Cc, n=1, boundary=bnd)
    assert same_bits(d_gpu, d), f"depth differs at {np.argwhere(d_gpu != d)[:5]}"
    assert same_bits(f_gpu, f)
    if N % 16:
        # the tail no work group covers kept its initial depth
        t = N // 16 * 16
        assert same_bits(d_gpu[t:, :], d0[t:, :]) and same_bits(d_gpu[:, t:], d0[:, t:])
