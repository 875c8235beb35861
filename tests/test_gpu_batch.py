"""The persistent batch engine (sfa_batch_*): many systems per launch.

Bars: FP64 batch solves match the oracle's ConjugateGradientSolverCSR_Rainbow
to the single-system bar (K within max(1, 2%), x within 1e-10 relative);
a solve's result is bit-identical alone and in any mix (its reductions never
see other slots); mixed-precision solves reach ||b - A x|| <= 1e-10 ||b||.
"""
import numpy as np
import pytest

import oracle
from simplefem_amd import capi
from simplefem_amd import systems as S

pytestmark = pytest.mark.gpu


def _systems():
    return [S.poisson2d_p1(3), S.poisson2d_p1(5), S.poisson3d_p1_kuhn(6), S.fd_laplacian_2d(12),
            S.poisson2d_p1(4), S.random_spd(300, 0.02, 5)]


def _add(batch, s, omega=1.0):
    col = oracle.rainbow(s.n, s.A, s.C)
    return batch.add_matrix(s.n, s.A, s.C, s.V, s.diagonal(), omega, *col), col


def _run_all(batch, slots, steps=0):
    print(f"[batch] run with {len(slots)} slots", flush=True)
    while batch.run(steps) > 0:
        pass
    done = set(batch.finished())
    assert done == set(slots), (done, slots)


def test_batch_fp64_matches_oracle():
    systems = _systems()
    with capi.Batch(mixed=False, max_slots=64) as batch:
        jobs = []
        for k, s in enumerate(systems):
            mid, col = _add(batch, s)
            for seed in (1, 2):
                b = S.rhs_uniform(s, 100 * k + seed)
                jobs.append((batch.submit(mid, b), s, b, col))
        _run_all(batch, [j[0] for j in jobs])
        for slot, s, b, col in jobs:
            x, info = batch.collect(slot, s.n)
            xo, ro, io = oracle.solve("rainbow", s.A, s.C, s.V, b, aux=s.diagonal(), colouring=col)
            Ko = io["iterations"]
            assert abs(info["iterations"] - Ko) <= max(1, int(0.02 * Ko)), (s.name, info, io)
            assert info["exit_reason"] == io["exit_reason"]
            assert np.abs(x - xo).max() <= 1e-10 * max(np.abs(xo).max(), 1e-300), s.name


def test_batch_result_is_the_same_alone_and_in_the_mix():
    systems = _systems()
    target = systems[2]
    bt = S.rhs_uniform(target, 77)
    with capi.Batch(max_slots=8) as alone:
        mid, _ = _add(alone, target)
        slot = alone.submit(mid, bt)
        _run_all(alone, [slot])
        x_alone, info_alone = alone.collect(slot, target.n)
    with capi.Batch(max_slots=64) as mix:
        slots = []
        for k, s in enumerate(systems):
            mid, _ = _add(mix, s)
            for seed in range(3):
                slots.append(mix.submit(mid, S.rhs_uniform(s, 5 + seed + 10 * k)))
        mid_t, _ = _add(mix, target)
        st = mix.submit(mid_t, bt)
        _run_all(mix, slots + [st])
        x_mix, info_mix = mix.collect(st, target.n)
    assert info_mix["iterations"] == info_alone["iterations"]
    assert np.array_equal(x_mix, x_alone)


def test_continuous_admission():
    """Solves submitted while others are running finish with the same result
    as when they run alone."""
    systems = _systems()
    with capi.Batch(max_slots=32) as batch:
        first = []
        for k, s in enumerate(systems[:3]):
            mid, _ = _add(batch, s)
            first.append((batch.submit(mid, S.rhs_uniform(s, k)), s, mid, k))
        batch.run(15)  # partly done
        late = []
        for k, s in enumerate(systems[3:]):
            mid, _ = _add(batch, s)
            late.append((batch.submit(mid, S.rhs_uniform(s, 50 + k)), s, mid, 50 + k))
        while batch.run(0) > 0:
            pass
        results = {}
        for slot, s, mid, seed in first + late:
            results[(s.name, seed)] = batch.collect(slot, s.n)
    for s, seed in [(systems[1], 1), (systems[4], 51)]:
        with capi.Batch(max_slots=4) as alone:
            mid, _ = _add(alone, s)
            slot = alone.submit(mid, S.rhs_uniform(s, seed))
            _run_all(alone, [slot])
            x, info = alone.collect(slot, s.n)
        xm, im = results[(s.name, seed)]
        assert np.array_equal(x, xm) and info["iterations"] == im["iterations"]


def _mixed_batch_run(inner_restarts, inner_rtol):
    systems = [S.poisson2d_p1(5), S.poisson3d_p1_kuhn(8), S.fd_laplacian_2d(20)]
    total = 0
    with capi.Batch(mixed=True, max_slots=16, inner_restarts=inner_restarts) as batch:
        jobs = []
        for k, s in enumerate(systems):
            mid, col = _add(batch, s)
            for seed in (1, 2, 3, 4):
                b = S.rhs_uniform(s, seed + 10 * k)
                jobs.append((batch.submit(mid, b, precision=1e-10, inner_rtol=inner_rtol), s, b))
        _run_all(batch, [j[0] for j in jobs])
        for slot, s, b in jobs:
            x, info = batch.collect(slot, s.n)
            assert info["success"] == 1, (s.name, info)
            res = b - oracle.csr_apply(s.n, s.A, s.C, s.V, x)
            res[s.diagonal() == 0] = 0.0
            # the refinement target is the reference's tolerance, max(precision*||b||, 2^-26)
            target = max(1e-10 * np.linalg.norm(b), 2.0 ** -26)
            assert np.linalg.norm(res) <= 1.0000001 * target, (s.name, info)
            assert info["outer"] >= 1
            total += info["iterations"]
    return total


@pytest.mark.parametrize("inner_restarts", [False, True], ids=["continued", "restarts"])
def test_batch_mixed_precision_reaches_fp64_residual(inner_restarts):
    _mixed_batch_run(inner_restarts, 1e-4)


def test_batch_mixed_continued_direction_needs_fewer_iterations():
    """Inner solves that continue the previous direction (the default) need
    fewer FP32 iterations in total than inner solves that restart."""
    cont = _mixed_batch_run(False, 1e-2)
    rest = _mixed_batch_run(True, 1e-2)
    print(f"[batch continued] inner iterations continued {cont} restarts {rest}", flush=True)
    assert cont < 0.8 * rest


def test_prepared_matrix_from_another_thread():
    """A plan built on a worker thread (sfa_batch_prepare_matrix) while the
    batch runs gives the same solve, bit for bit, as sfa_batch_add_matrix."""
    from concurrent.futures import ThreadPoolExecutor

    systems = _systems()[:4]
    cols = [oracle.rainbow(s.n, s.A, s.C) for s in systems]
    with capi.Batch(max_slots=16) as batch, ThreadPoolExecutor(2) as pool:
        futs = [pool.submit(capi.PreparedMatrix, s.n, s.A, s.C, s.V, s.diagonal(), 1.0, *c)
                for s, c in zip(systems, cols)]
        direct = []
        for k, s in enumerate(systems):  # the batch runs while the plans are built
            mid, _ = _add(batch, s)
            direct.append(batch.submit(mid, S.rhs_uniform(s, 40 + k)))
            batch.run(32)
        prepared = []
        for k, (s, f) in enumerate(zip(systems, futs)):
            p = f.result()
            mid = batch.add_prepared(p)
            p.close()
            prepared.append(batch.submit(mid, S.rhs_uniform(s, 40 + k)))
        _run_all(batch, direct + prepared)
        for s, a, b in zip(systems, direct, prepared):
            xa, ia = batch.collect(a, s.n)
            xb, ib = batch.collect(b, s.n)
            assert np.array_equal(xa, xb) and ia["iterations"] == ib["iterations"], s.name


@pytest.mark.parametrize("mixed", [False, True])
def test_right_hand_sides_riding_together_match_alone(mixed):
    """Adjacent slots of one matrix in the same phase share each unit (the
    matrix is read once for all of them, up to four per unit); each one's x
    and iteration count are bit-identical to a solve alone.  Six right-hand
    sides per matrix: groups of four and two, which split as solves finish."""
    systems = [S.poisson3d_p1_kuhn(7), S.poisson2d_p1(5), S.random_spd(250, 0.05, 3)]
    jobs = []
    with capi.Batch(mixed=mixed, max_slots=32) as batch:
        for k, s in enumerate(systems):
            mid, _ = _add(batch, s)
            for seed in range(6):
                b = S.rhs_uniform(s, 300 + 10 * k + seed)
                jobs.append((batch.submit(mid, b), s, b))
        _run_all(batch, [j[0] for j in jobs])
        together = [batch.collect(slot, s.n) for slot, s, _ in jobs]
    for (slot, s, b), (x_t, info_t) in zip(jobs, together):
        with capi.Batch(mixed=mixed, max_slots=2) as alone:
            mid, _ = _add(alone, s)
            sl = alone.submit(mid, b)
            _run_all(alone, [sl])
            x_a, info_a = alone.collect(sl, s.n)
        assert info_t["iterations"] == info_a["iterations"], (s.name, info_t, info_a)
        assert np.array_equal(x_t, x_a), s.name


@pytest.mark.parametrize("mixed", [False, True])
def test_interleaved_groups_match_separate_slots(mixed):
    """Right-hand sides submitted as one interleaved group
    (sfa_batch_submit_group: vectors stored row-interleaved, a complete group
    of four gathering with one vector load) give the x and iteration counts of
    the same right-hand sides submitted one slot at a time, bit for bit.
    Groups of 4, 3, 2 and 1 over matrices whose solves converge at different
    iteration counts, so groups split as their solves finish; initial guesses
    nonzero for one group."""
    systems = [S.poisson3d_p1_kuhn(7), S.poisson2d_p1(5), S.random_spd(250, 0.05, 3)]
    sizes = [4, 3, 2, 1]
    with capi.Batch(mixed=mixed, max_slots=64) as grouped, \
            capi.Batch(mixed=mixed, max_slots=64) as separate:
        jobs = []
        for k, s in enumerate(systems):
            mg, _ = _add(grouped, s)
            ms, _ = _add(separate, s)
            for gi, g in enumerate(sizes):
                bs = [S.rhs_uniform(s, 500 + 40 * k + 10 * gi + i) for i in range(g)]
                x0s = None
                if gi == 1:
                    rng = np.random.default_rng(9 + k)
                    x0s = [rng.uniform(-1, 1, s.n) for _ in range(g)]
                gslots = grouped.submit_group(mg, bs, x0s)
                assert gslots == list(range(gslots[0], gslots[0] + g))
                sslots = [separate.submit(ms, b, None if x0s is None else x0s[i])
                          for i, b in enumerate(bs)]
                jobs += [(gs, ss, s) for gs, ss in zip(gslots, sslots)]
        print(f"case: interleaved groups {'mixed' if mixed else 'fp64'}, {len(jobs)} solves",
              flush=True)
        _run_all(grouped, [j[0] for j in jobs])
        _run_all(separate, [j[1] for j in jobs])
        iters = set()
        for gs, ss, s in jobs:
            xg, ig = grouped.collect(gs, s.n)
            xs, is_ = separate.collect(ss, s.n)
            assert ig["iterations"] == is_["iterations"], (s.name, ig, is_)
            assert ig["exit_reason"] == is_["exit_reason"] and ig["success"] == 1
            assert np.array_equal(xg, xs), s.name
            iters.add((s.name, ig["iterations"]))
        # the groups did split: one matrix's solves end at different counts
        assert len(iters) > len(systems), iters
        assert grouped.memory() > 0


@pytest.mark.parametrize("mixed", [False, True])
def test_group_lanes_run_on_after_lane_zero_slot_is_reused(mixed):
    """A group's lane 0 finishes first (a right-hand side of 1e-12 scale is
    within the 2^-26 floor at once) and is collected; its slot is then taken
    by a single solve of a larger matrix while lanes 1..3 still run.  The
    lanes keep using their own group's vectors: every result equals the
    same solve in separate slots, bit for bit."""
    sg, sb = S.poisson3d_p1_kuhn(6), S.poisson2d_p1(6)
    bs = [1e-12 * S.rhs_uniform(sg, 700)] + [S.rhs_uniform(sg, 701 + i) for i in range(3)]
    bb = S.rhs_uniform(sb, 710)
    with capi.Batch(mixed=mixed, max_slots=8) as batch:
        mg, _ = _add(batch, sg)
        mb, _ = _add(batch, sb)
        slots = batch.submit_group(mg, bs)
        print(f"case: lane 0 reuse {'mixed' if mixed else 'fp64'}, group slots {slots}", flush=True)
        for _ in range(50):
            batch.run(32)
            done = set(batch.finished())
            if slots[0] in done:
                break
        assert slots[0] in done and not (set(slots[1:]) & done), done
        x0, i0 = batch.collect(slots[0], sg.n)
        single = batch.submit(mb, bb)
        assert single == slots[0]  # the freed slot of lane 0
        _run_all(batch, slots[1:] + [single])
        got = [batch.collect(s, sg.n) for s in slots[1:]] + [batch.collect(single, sb.n)]
    want = []
    with capi.Batch(mixed=mixed, max_slots=8) as sep:
        ms, _ = _add(sep, sg)
        mbs, _ = _add(sep, sb)
        ss = [sep.submit(ms, b) for b in bs] + [sep.submit(mbs, bb)]
        _run_all(sep, ss)
        want = [sep.collect(s, sg.n) for s in ss[:4]] + [sep.collect(ss[4], sb.n)]
    assert np.array_equal(x0, want[0][0]) and i0["iterations"] == want[0][1]["iterations"]
    for (xg, ig), (xw, iw) in zip(got, want[1:]):
        assert ig["iterations"] == iw["iterations"] and ig["exit_reason"] == iw["exit_reason"]
        assert np.array_equal(xg, xw)
