"""The HIP step (through the C ABI) against the reference's own shaders.

Same fixtures as test_golden.py (unmodified reference GLSL on Mesa llvmpipe);
bar: bit-identical depth, flux and rg16f FlowMap bits, for every tile kernel
variant and steps-per-launch setting used in production.
"""
import json
from pathlib import Path

import numpy as np
from bitwise import same_bits
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"
MANIFEST = json.loads((GOLDEN / "manifest.json").read_text())
CASES = sorted(MANIFEST["cases"])
LAUNCHES = [(0, 0), (1, 1), (2, 1), (2, 8), (3, 4), (3, 5), (4, 2), (5, 3), (5, 8), (6, 2), (6, 8),
            (7, 3), (7, 6), (7, 8), (8, 3), (8, 5), (8, 8)]


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("kernel,k", LAUNCHES)
def test_hip_matches_reference_shaders(gpu, name, kernel, k):
    case = MANIFEST["cases"][name]
    z = np.load(GOLDEN / f"{name}.npz")
    H, W = case["shape"]
    boundary = {"closed": gpu.BOUNDARY_CLOSED, "zero": gpu.BOUNDARY_ZERO,
                "dispatch": gpu.BOUNDARY_REFERENCE_DISPATCH}[case["boundary"]]
    p = gpu.params_raw(case["F_value"], case["K_value"], case["C_value"], boundary)
    for n in case["steps"]:
        with gpu.Simulation(W, H, p) as s:
            if kernel:
                s.set_launch(kernel, k)
            s.upload(z["b"], z["d0"], z["f0"] if "f0" in z.files else None)
            if case["rain"]:
                s.set_rain(*case["rain"])
            s.step(n)
            out = s.download(flux=True)
            ref = s.download_reference_layout()
        assert same_bits(out["depth"], z[f"d_{n}"]), f"{name}: depth after {n} steps"
        assert same_bits(out["flux"], z[f"f_{n}"]), f"{name}: flux after {n} steps"
        assert same_bits(ref["flowmap_rg16f"], z[f"v16_{n}"]), f"{name}: FlowMap"


WINDOWS = sorted(MANIFEST.get("window_cases", {}))


@pytest.mark.parametrize("name", WINDOWS)
@pytest.mark.parametrize("kernel,k", [(0, 0), (3, 4), (2, 8), (8, 4), (8, 5), (8, 6)])
def test_hip_matches_reference_window(gpu, name, kernel, k):
    """The headline configuration (4096^2 fBm terrain + 0.5 m sheet, closed
    walls) after 64 steps on the default launch (and two others): the centre
    256^2 window equals the reference shaders' result on llvmpipe."""
    case = MANIFEST["window_cases"][name]
    z = np.load(GOLDEN / f"{name}.npz")
    N = case["shape"][0]
    (n,) = case["steps"]
    y0, x0, wh, ww = case["window"]
    b = gpu.generate_fbm_terrain(N, N, seed=2, octaves=6, base_frequency=4, persistence=0.43,
                                 height_scale=100.0)
    p = gpu.params_raw(case["F_value"], case["K_value"], case["C_value"], gpu.BOUNDARY_CLOSED)
    print(f"window {name}: launch {kernel}/{k}", flush=True)
    with gpu.Simulation(N, N, p) as s:
        if kernel:
            s.set_launch(kernel, k)
        s.upload(b, np.full((N, N), 0.5, np.float32))
        s.step(n)
        out = s.download(flux=True)
        ref = s.download_reference_layout()
    w = (slice(y0, y0 + wh), slice(x0, x0 + ww))
    assert same_bits(b[w], z["b_window"])
    assert same_bits(out["depth"][w], z[f"d_{n}"]), "depth window differs"
    assert same_bits(out["flux"][w], z[f"f_{n}"]), "flux window differs"
    assert same_bits(ref["flowmap_rg16f"][w], z[f"v16_{n}"]), "FlowMap window differs"
