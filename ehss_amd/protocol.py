"""The publication sampling protocol for one CCS estimate.

* Three independently scrambled Sobol blocks with seeds s, s + 10007 and
  s + 20014 (mod 2^32), each of 2^power points, pooled with equal weights, so
  the pooled estimate is the mean of the three block estimates
  (manuscript/source/supporting_information.md:21, manuscript.md:33).
* The blocks are traced at collision cap 512 and extended to 1,024, 2,048 and
  4,096 while the sampled unresolved-path bound exceeds 0.05 % of the computed
  CCS; every attempt is kept (manuscript.md:61).  The reference retraces the
  same rays at each cap.  Here only the unresolved rays are traced further,
  from their saved state (ehss_solve_extend): the trace is deterministic, so
  the result is bit-identical to the retrace.

The three blocks run as one multi-replicate launch, extended together.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

BLOCK_STRIDE = 10007
CAPS = (512, 1024, 2048, 4096)
TAIL_FRACTION = 5e-4


def block_seeds(seed: int, blocks: int = 3) -> list[int]:
    return [(int(seed) + BLOCK_STRIDE * k) % 2 ** 32 for k in range(blocks)]


@dataclass
class Attempt:
    cap: int
    omega: float            # pooled CCS estimate (A^2)
    pa: float               # pooled PA estimate (A^2)
    tail_bound: float       # pooled sampled bound on the unresolved contribution (A^2)
    unresolved: int
    block_omegas: list = field(default_factory=list)


@dataclass
class ProtocolResult:
    seed: int
    power: int
    attempts: list
    accepted: bool          # the final attempt's bound is within the fraction

    @property
    def final(self) -> Attempt:
        return self.attempts[-1]

    @property
    def omega(self) -> float:
        return self.final.omega


def publication_estimate(scene, seed: int, power: int, caps=CAPS, tail_fraction: float = TAIL_FRACTION,
                         precision: str = "fp32-fixup", pa: bool = True, scramble=None) -> ProtocolResult:
    """One pooled three-block estimate with cap extension on a device scene
    (anything with GPUScene.solve_many)."""
    if scramble is None:
        from .engine import scramble
    caps = [int(c) for c in caps]
    if not caps or any(b <= a for a, b in zip(caps, caps[1:])):
        raise ValueError("caps must be a nonempty increasing sequence")
    sobols = [scramble(s) for s in block_seeds(seed)]
    attempts = []
    accepted = False
    for k, cap in enumerate(caps):
        if k == 0:
            res = scene.solve_many(sobols, power, cap, precision=precision, pa=pa, continuable=True)
        else:
            res = scene.extend(cap, replicates=len(sobols))
        om = [r["omega"] for r in res]
        att = Attempt(cap=cap, omega=float(np.mean(om)), pa=float(np.mean([r["pa"] for r in res])),
                      tail_bound=float(np.mean([r["tail_bound"] for r in res])),
                      unresolved=int(sum(r["unresolved"] for r in res)), block_omegas=om)
        attempts.append(att)
        accepted = att.tail_bound <= tail_fraction * att.omega
        if accepted:
            break
    return ProtocolResult(seed=int(seed), power=int(power), attempts=attempts, accepted=accepted)
