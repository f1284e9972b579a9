"""The C++ host mirror of the reference's Terrain class (include/tws_terrain.hpp),
driven by tests/cpp/test_terrain.cpp (built by __graft_entry__.build())."""
import subprocess
from pathlib import Path

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
def test_cpp_mirror_matches_c_abi(gpu):
    if not EXE.exists():
        pytest.fail("tests/cpp/_build/test_terrain missing: run __graft_entry__.build()")
    res = subprocess.run([str(EXE)], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0 and res.stdout.strip().endswith("OK"), res.stdout + res.stderr
