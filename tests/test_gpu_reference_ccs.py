"""Full-scale statistical agreement with the reference's own FP64 CCS values
(data/benchmark/reference_summary.json: 8 independent 2^20-ray sphere-boundary
estimates per structure, heavy atoms, C/N/O/S radii + 1.4 A probe, caps up to
4096; compared by a z-test on the two standard errors), for all 18 evaluation inputs shipped with the reference
(data/inputs/evaluation18, committed gzip-compressed under
tests/golden/evaluation18), 92 to 49,676 atoms."""
import gzip
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


@pytest.mark.parametrize("pdb", sorted(REF, key=lambda k: REF[k]["atoms"]))
def test_ccs_agrees_with_reference_fp64_estimate(pdb, tmp_path):
    xyz = tmp_path / f"{pdb}.xyz"
    xyz.write_bytes(gzip.decompress((GOLDEN / "evaluation18" / f"{pdb}.xyz.gz").read_bytes()))
    spheres, _ = prepare_frame(read_xyz(xyz)[0], radius_table(), 1.4)
    ref = REF[pdb]
    assert len(spheres.radii) == ref["atoms"]
    # 32 replicates of 2^20 rays in one launch: 8, as the reference used, give a
    # poor estimate of our own standard error (one structure drew z = -3.4 at 8
    # and converged to z = -0.6 at 64).
    with GPUScene(spheres.centers, spheres.radii) as sc:
        out = sc.solve_many([scramble(replicate_seed(20240, r)) for r in range(32)], 20, 4096,
                            precision="fp32-fixup", pa=False)
    vals = [res["omega"] for res in out]
    tails = [res["tail_bound"] for res in out]
    mean, sd = float(np.mean(vals)), float(np.std(vals, ddof=1))
    se_ours = sd / np.sqrt(len(vals))
    se_ref = ref["half95"] / 2.365  # t(0.975, 7)
    z = (mean - ref["mean"]) / np.hypot(se_ours, se_ref)
    tail_pct = 100 * max(tails) / mean
    say(f"{pdb} ({ref['atoms']} atoms): ours {mean:.5f} +- {se_ours:.5f}, reference {ref['mean']:.5f} +- "
        f"{se_ref:.5f}, z = {z:.2f}; max tail {tail_pct:.5f} % (reference {ref['max_tail_percent']:.5f} %)")
    assert abs(z) < 4.5
    assert tail_pct <= 0.05  # the protocol's acceptance bound on the unresolved tail
