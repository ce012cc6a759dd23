"""Writes tests/golden/sobol_scipy.npz: black-box vectors of SciPy's scrambled
Sobol generator (scipy.stats.qmc.Sobol, the third-party source of the reference's
incidence, called at ehss_state.py:44 as Sobol(4, scramble=True, seed=seed)
.random_base2(power)).

Recorded with SciPy 1.15.3 / NumPy 2.2.6.  Stored as integers q = u * 2^30.
Run: python tests/golden/make_sobol_golden.py
"""
from pathlib import Path

import numpy as np
import scipy
from scipy.stats import qmc

SEEDS = [0, 1, 2, 98, 881, 104729, 209458, 190003, 4294967295]
FULL = {0, 881, 104729, 4294967295}  # seeds stored with 2^10 points (others 2^6)
OUT = Path(__file__).resolve().parent / "sobol_scipy.npz"


def q_of(u):
    q = u * 2.0 ** 30
    assert np.array_equal(q, np.round(q))
    return q.astype(np.uint32)


def main():
    data = {"scipy_version": np.array(scipy.__version__), "numpy_version": np.array(np.__version__),
            "seeds": np.array(SEEDS, dtype=np.uint64)}
    for dim in (2, 4, 5):
        for seed in SEEDS:
            eng = qmc.Sobol(dim, scramble=True, seed=seed)
            data[f"sv_d{dim}_s{seed}"] = eng._sv.astype(np.uint32)
            data[f"shift_d{dim}_s{seed}"] = eng._shift.astype(np.uint32)
            m = 10 if (dim == 4 and seed in FULL) else 6
            data[f"q_d{dim}_s{seed}_m{m}"] = q_of(qmc.Sobol(dim, scramble=True, seed=seed).random_base2(m))
    rng = np.random.default_rng(7)
    for seed in (0, 1):
        for m in (20, 22):
            u = qmc.Sobol(4, scramble=True, seed=seed).random_base2(m)
            idx = np.sort(rng.choice(2 ** m, size=256, replace=False))
            data[f"idx_d4_s{seed}_m{m}"] = idx.astype(np.uint32)
            data[f"qidx_d4_s{seed}_m{m}"] = q_of(u[idx])
    np.savez_compressed(OUT, **data)
    print(f"wrote {OUT} ({OUT.stat().st_size} bytes)")


if __name__ == "__main__":
    main()
