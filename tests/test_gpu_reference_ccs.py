"""Full-scale statistical agreement with the reference's own FP64 CCS values
(data/benchmark/reference_summary.json: 8 independent 2^20-ray sphere-boundary
estimates per structure, heavy atoms, C/N/O/S radii + 1.4 A probe), for the
evaluation inputs shipped with the reference (data/inputs/evaluation18)."""
import json
from pathlib import Path

import numpy as np
import pytest

from ehss_amd.ccs_cli import prepare_frame, radius_table, read_xyz
from ehss_amd.engine import GPUScene, replicate_seed, scramble
from tests.gpu_helpers import say

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"
REF = json.loads((GOLDEN / "reference_ccs.json").read_text())["structures"]


@pytest.mark.parametrize("pdb", ["1AL1", "1A11", "1VII"])
def test_ccs_agrees_with_reference_fp64_estimate(pdb):
    spheres, _ = prepare_frame(read_xyz(GOLDEN / "examples" / f"{pdb}.xyz")[0], radius_table(), 1.4)
    ref = REF[pdb]
    assert len(spheres.radii) == ref["atoms"]
    vals = []
    with GPUScene(spheres.centers, spheres.radii) as sc:
        for r in range(8):
            res = sc.solve(scramble(replicate_seed(20240, r)), 20, 4096, precision="fp32-fixup", pa=False)
            assert res["unresolved"] == 0
            vals.append(res["omega"])
    mean, sd = float(np.mean(vals)), float(np.std(vals, ddof=1))
    se_ours = sd / np.sqrt(len(vals))
    se_ref = ref["half95"] / 2.365  # t(0.975, 7)
    z = (mean - ref["mean"]) / np.hypot(se_ours, se_ref)
    say(f"{pdb}: ours {mean:.5f} +- {se_ours:.5f}, reference {ref['mean']:.5f} +- {se_ref:.5f}, z = {z:.2f}")
    assert abs(z) < 4.5
    assert abs(mean - ref["mean"]) / ref["mean"] < 5e-4
