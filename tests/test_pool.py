"""Host logic of the batch's device pool (terrainwatersim_amd/csrc/tws_pool.hpp),
exercised on the CPU through a small driver compiled with g++."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent

DRIVER = r"""
#include "tws_pool.hpp"
#include <cstdio>
#include <random>
#include <vector>
struct B { size_t off, bytes; };
int main() {
    tws::DevicePool p;
    const size_t cap = 1 << 24;
    p.reset(reinterpret_cast<char*>(4096), cap);
    std::mt19937_64 rng(7);
    std::vector<B> live;
    for (int it = 0; it < 20000; ++it) {
        if (live.empty() || rng() % 3) {
            size_t want = 1 + rng() % (1 << 18);
            size_t off = p.alloc(want);
            if (off == SIZE_MAX) continue;
            if (off % tws::DevicePool::kAlign) { printf("misaligned\n"); return 1; }
            if (off + tws::DevicePool::round(want) > cap) { printf("overflow\n"); return 1; }
            for (auto& b : live)
                if (off < b.off + tws::DevicePool::round(b.bytes) &&
                    b.off < off + tws::DevicePool::round(want)) { printf("overlap\n"); return 1; }
            live.push_back({off, want});
        } else {
            size_t i = rng() % live.size();
            p.release(live[i].off, live[i].bytes);
            live.erase(live.begin() + i);
        }
        size_t used = 0;
        for (auto& b : live) used += tws::DevicePool::round(b.bytes);
        if (used != p.used()) { printf("accounting\n"); return 1; }
    }
    for (auto& b : live) p.release(b.off, b.bytes);
    if (p.used() != 0 || p.largest_free() != cap) { printf("no coalesce\n"); return 1; }
    if (p.alloc(cap) != 0 || p.alloc(1) != SIZE_MAX) { printf("full\n"); return 1; }
    printf("OK\n");
    return 0;
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_pool_invariants(tmp_path):
    src = tmp_path / "pool_driver.cpp"
    src.write_text(DRIVER)
    exe = tmp_path / "pool_driver"
    subprocess.run(["g++", "-O1", "-std=c++17", "-fsanitize=address,undefined",
                    f"-I{ROOT / 'terrainwatersim_amd' / 'csrc'}", str(src), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0 and out.stdout.strip() == "OK", out.stdout + out.stderr
