"""The HIP step (through the C ABI) against the reference's own shaders.

Same fixtures as test_golden.py (unmodified reference GLSL on Mesa llvmpipe);
bar: bit-identical depth, flux and rg16f FlowMap bits, for every tile kernel
variant and steps-per-launch setting used in production.
"""
import json
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"
MANIFEST = json.loads((GOLDEN / "manifest.json").read_text())
CASES = sorted(MANIFEST["cases"])
LAUNCHES = [(0, 0), (1, 1), (2, 1), (2, 8), (3, 4), (3, 5), (4, 2), (5, 3), (5, 8), (6, 2), (6, 8),
            (7, 3), (7, 6), (7, 8)]


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("kernel,k", LAUNCHES)
def test_hip_matches_reference_shaders(gpu, name, kernel, k):
    case = MANIFEST["cases"][name]
    z = np.load(GOLDEN / f"{name}.npz")
    H, W = case["shape"]
    boundary = gpu.BOUNDARY_CLOSED if case["boundary"] == "closed" else gpu.BOUNDARY_ZERO
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
        assert np.array_equal(out["depth"], z[f"d_{n}"]), f"{name}: depth after {n} steps"
        assert np.array_equal(out["flux"], z[f"f_{n}"]), f"{name}: flux after {n} steps"
        assert np.array_equal(ref["flowmap_rg16f"], z[f"v16_{n}"]), f"{name}: FlowMap"
