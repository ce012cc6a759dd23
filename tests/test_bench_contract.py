"""The benchmark's fixed parts, without a GPU: the roofline figures it reads from
the committed rocprofv3 summaries parse, and a run asked for several GPUs exits
with a message."""
import subprocess
import sys
from pathlib import Path

import bench

ROOT = Path(__file__).resolve().parent.parent


def test_committed_profiles_give_the_roofline_figures():
    traffic = bench.read_traffic()
    valu = bench.read_valu_issue()
    assert traffic is not None and traffic > 16 * (1 << 20) * 12  # at least the records of a launch
    assert valu is not None and 0.0 < valu <= 1.0


def test_more_than_one_gpu_is_refused_with_a_message():
    out = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode != 0
    assert "one GPU" in (out.stdout + out.stderr) or "1 GPU" in (out.stdout + out.stderr)
