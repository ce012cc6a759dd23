#!/usr/bin/env python3
"""Run with: python ccs.py molecule.xyz --pa-method bvh --output results/molecule"""
from pathlib import Path
import os
import sys

if __name__ == "__main__":
    for name in ("OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS", "OMP_NUM_THREADS"):
        os.environ.setdefault(name, "1")
    sys.path.insert(0, str(Path(__file__).resolve().parent))
    from ehss_amd.ccs_cli import main
    raise SystemExit(main())
