"""The library's block cache (csrc/pool.cpp): Systems created, solved and
destroyed from several host threads at once, each on its own stream, take
blocks the others freed.  Every owner waits for its own stream before it
releases, so every solve must equal, bit for bit, the same solve run alone."""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from simplefem_amd import capi
from simplefem_amd import systems as S

pytestmark = pytest.mark.gpu


def _solve(s, b, mixed):
    rb = capi.Rainbow(s.n, s.A, s.C)
    d = capi.get_diagonal(s.n, s.A, s.C, s.V)
    with capi.System(capi.METHOD_RAINBOW, s.n, s.A, s.C, s.V, d, 1.0, rb.num_colors, rb.F, rb.B,
                     rb.R, fp32_inner=mixed, borrow_csr=True) as sysd:
        x = np.zeros(s.n)
        if mixed:
            assert sysd.solve_mixed(x, b, 1e-10)["success"]
        else:
            info = {}
            sysd.solve(x, b, np.zeros(s.n), 1e-10, info=info)
            assert info["success"]
    return x


def test_concurrent_systems_share_the_block_cache():
    shapes = [(2, 6, 1), (3, 5, 1), (2, 20, 2), (3, 4, 2), (2, 12, 3), (2, 40, 1)]  # dim, n, r
    jobs = []
    for k, (dim, n, r) in enumerate(shapes * 3):
        s = S.lagrange_pr(dim, n, r)
        jobs.append((s, S.rhs_uniform(s, 100 + k), k % 2 == 1))
    print(f"case: {len(jobs)} systems, {sum(j[0].n for j in jobs)} rows in all", flush=True)
    alone = [_solve(*j) for j in jobs]
    with ThreadPoolExecutor(6) as pool:
        together = list(pool.map(lambda j: _solve(*j), jobs))
    for k, (a, t) in enumerate(zip(alone, together)):
        assert np.array_equal(a, t), k
