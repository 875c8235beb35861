"""Workloads of the mixed-precision, batched and served forms.

* ``batched_systems``  256 independent 2D Poisson systems: per system a seeded
  refinement level L in 4..8 (UnitSquare2D_simple refined L times) and
  Lagrange degree r in 1..3, (r*2^L+1)^2 unknowns, each with 4 seeded
  U[-1,1) right-hand sides.  The matrices are the reference's: its degree-r
  stiffness in FEEC's Sullivan basis on its own refined mesh
  (``systems.sullivan_pr_2d``).
* ``served_requests``  a Poisson arrival stream; each request draws the
  dimension (2 or 3), a log-uniform unknown count in [lo, hi] and a degree
  1..3, and carries one seeded right-hand side.  Requests of equal shape share
  their generated matrix (the generator is deterministic).  Their meshes have
  any number of cells per axis, which the reference's refinement (2^L only)
  cannot produce, so they use the package's own nodal Lagrange assembly
  (``systems.lagrange_pr``).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import systems as S


def _rng(seed):
    return np.random.default_rng(seed)


@dataclass
class Job:
    system: S.CSRSystem
    rhs: list = field(default_factory=list)
    arrival: float = 0.0


def batched_systems(count=256, nrhs=4, seed=2026, levels=(4, 8), degrees=(1, 3)):
    rng = _rng(seed)
    jobs = []
    for k in range(count):
        L = int(rng.integers(levels[0], levels[1] + 1))
        r = int(rng.integers(degrees[0], degrees[1] + 1))
        s = S.sullivan_pr_2d(L, r)
        rhs = [S.rhs_uniform(s, seed * 100003 + k * 17 + j) for j in range(nrhs)]
        jobs.append(Job(s, rhs))
    return jobs


def served_requests(count=64, rate=50.0, seed=7, dofs=(1e3, 2e6), degrees=(1, 3),
                    progress=None):
    """count requests with exponential inter-arrival times (rate per second).
    progress(k, job), if given, is called after request k is generated."""
    rng = _rng(seed)
    t = 0.0
    jobs = []
    for k in range(count):
        t += float(rng.exponential(1.0 / rate))
        dim = int(rng.integers(2, 4))
        r = int(rng.integers(degrees[0], degrees[1] + 1))
        target = float(np.exp(rng.uniform(np.log(dofs[0]), np.log(dofs[1]))))
        n = max(2, int(round((target ** (1.0 / dim) - 1.0) / r)))
        s = S.lagrange_pr(dim, n, r)
        jobs.append(Job(s, [S.rhs_uniform(s, seed * 7919 + k)], arrival=t))
        if progress is not None:
            progress(k, jobs[-1])
    return jobs
