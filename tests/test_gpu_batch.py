"""Batched stepping (vcp_batch_*) through the C ABI.

Many Taylor-Green cases of uneven polynomial degree (P = 2..6, so every
(K, M) group of the library), element count, Mach and Reynolds numbers and
phase shifts share one launch per kernel and group. Each case carries its own
time step and synchronization-step clipping on the device. The bars:
- each case after k RK steps in the batch is bit-identical to the same case
  run alone in a batch of one with the same volume workgroup size (the same
  kernels, the same arithmetic in the same order), and within relative L2
  1e-12 of the C oracle;
- against the single-case API (vcp_sync_step, whose kernels are separate
  instantiations of the same source, compiled without the case-table
  indirection) within relative L2 1e-14 (rounding; observed 0 on most cases);
- the graph-replayed batch is bit-identical to the launch-by-launch batch;
- a case whose state makes the time step fail (a^2 negative or not a number)
  is deactivated while the others continue unchanged.
"""
import math

import numpy as np
import pytest

from vcp3d_open_amd.cases import OpenBoxCase, TGVCase

pytestmark = pytest.mark.gpu

STEPS_DONE_ = 1     # status word of vcp_batch_status: RK steps done


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def lib():
    from vcp3d_open_amd.solver import load_library
    return load_library()


def seeded_cases(seed=7, count=7):
    """TGV cases drawn like the batched workload, at sizes the oracle runs in
    seconds: P uniform in {2..6} (each degree at least once), n in {2, 3}
    (Cartesian) or 4 (curvilinear: on coarser grids the perturbed nodes no
    longer pass the LGL spacing check of the grid reader), M in [0.05, 0.4],
    Re log-uniform in [400, 3200], phases in [0, 2 pi)^3; every other case
    curvilinear."""
    rng = np.random.default_rng(seed)
    ps = [2, 3, 4, 5, 6] + list(rng.integers(2, 7, size=count - 5))
    out = []
    for i, p in enumerate(ps):
        n = int(rng.integers(2, 4))
        out.append(dict(n=4 if i % 2 == 1 else n, p=int(p), mach=float(rng.uniform(0.05, 0.4)),
                        reynolds=float(math.exp(rng.uniform(math.log(400), math.log(3200)))),
                        phase=tuple(float(x) for x in rng.uniform(0, 2 * math.pi, 3)),
                        curvilinear=(i % 2 == 1)))
    return out


def with_sync_steps(kw, factor):
    """The case with its synchronization time step set to factor times its
    smallest element time step (factor > 1: several RK steps per call)."""
    from vcp3d_open_amd.solver import SolverClass
    case = TGVCase(**kw)
    probe = SolverClass(case)
    probe.ReadRestartSolution(case.initial_conservative())
    dt_min = probe.time_steps().min() / probe.reference()["u_ref"]
    probe.close()
    return TGVCase(**kw, dt_synchr=factor * dt_min)


def run_alone(case, n_sync, nt=0):
    from vcp3d_open_amd.solver import SolverClass
    g = SolverClass(case)
    if nt:
        g.set_tile(nt)
    g.ReadRestartSolution(case.initial_conservative())
    steps = 0
    for _ in range(n_sync):
        g.ExecuteTimeSynchronizationStep()
        steps += g.last_rk_steps
    return g, steps


def run_batch_of_one(case, n_steps, nt=0, damping=None):
    """The case stepped n_steps RK steps alone in a batch of one."""
    from vcp3d_open_amd.solver import Batch, SolverClass
    g = SolverClass(case)
    g.ReadRestartSolution(case.initial_conservative())
    if damping is not None:
        g.SetDampingState(damping)
    K, M = g.K, g.M
    b = Batch(tiles={(K, M): nt} if nt else None)
    b.add(g, n_steps)
    b.step(n_steps)
    b.wait()
    b.close()
    return g


def test_batch_matches_alone_and_oracle(lib, oracle_lib):
    from vcp3d_open_amd.solver import Batch, SolverClass, ACTIVE, STEPS_DONE, SYNC_DONE
    cases = [with_sync_steps(kw, f) for kw, f in zip(seeded_cases(), [0.7, 1.6, 0.9, 2.3, 0.8, 1.3, 3.1])]
    n_sync = 2
    alone = [run_alone(c, n_sync) for c in cases]
    batch = Batch()
    members = []
    for c, (_, steps) in zip(cases, alone):
        g = SolverClass(c)
        g.ReadRestartSolution(c.initial_conservative())
        batch.add(g, steps)
        members.append(g)
    groups = batch.groups()
    assert len(groups) == 5, groups
    print(f"[case] batch of {len(cases)} TGV cases, groups {[(x['K'], x['M'], x['cases']) for x in groups]}",
          flush=True)
    batch.step(max(s for _, s in alone))
    st = batch.status()
    for i, (c, (a, steps), g) in enumerate(zip(cases, alone, members)):
        assert st[g][STEPS_DONE] == steps and st[g][SYNC_DONE] == n_sync and st[g][ACTIVE] == 0, (i, st[g])
        one = run_batch_of_one(c, steps)
        assert np.array_equal(g.solution(), one.solution()), f"case {i}: state differs from the case alone"
        assert np.array_equal(g.residual_array(), one.residual_array()), f"case {i}: residual differs"
        assert rel(g.solution(), a.solution()) < 1e-14, (i, rel(g.solution(), a.solution()))
        o = oracle_lib.Oracle(c)
        o.set_restart(c.initial_conservative())
        for _ in range(n_sync):
            o.sync_step()
        assert rel(g.solution(), o.modal()) < 1e-12, (i, rel(g.solution(), o.modal()))
    # members leave the batch and continue alone from where they are
    for g in members:
        batch.remove(g)
    for g, (a, _) in zip(members, alone):
        g.ExecuteTimeSynchronizationStep()
        a.ExecuteTimeSynchronizationStep()
        assert rel(g.solution(), a.solution()) < 1e-14


def test_graph_batch_matches_eager_batch(lib):
    from vcp3d_open_amd.solver import Batch, SolverClass
    cases = [TGVCase(**kw) for kw in seeded_cases(seed=11, count=6)]
    print("[case] graph batch vs eager batch, 6 cases", flush=True)
    eager, graph = Batch(), Batch(use_graph=True)
    pairs = []
    for c in cases:
        ge, gg = SolverClass(c), SolverClass(c)
        for g in (ge, gg):
            g.ReadRestartSolution(c.initial_conservative())
        eager.add(ge, 0)
        graph.add(gg, 0)
        pairs.append((ge, gg))
    for rnd in range(3):
        eager.step(2)
        graph.step(2)
        eager.wait()
        graph.wait()
        for ge, gg in pairs:
            assert np.array_equal(gg.solution(), ge.solution()), f"round {rnd}"
            assert np.array_equal(gg.residual_array(), ge.residual_array()), f"round {rnd}"
        if rnd == 0:
            # a membership change recaptures the graph
            eager.remove(pairs[2][0])
            graph.remove(pairs[2][1])
            pairs.pop(2)


def test_open_box_in_batch_matches_alone(lib):
    from vcp3d_open_amd.solver import Batch, SolverClass
    box = OpenBoxCase(n=(6, 2, 2), p=3, sponge_elems=2)
    tgv = TGVCase(n=3, p=3)
    print("[case] open box 6x2x2 P=3 + TGV n3 P=3 in one batch", flush=True)

    def make(c):
        g = SolverClass(c)
        g.ReadRestartSolution(c.initial_conservative())
        if c is box:
            g.SetDampingState(box.damping_primitive())
        return g

    a_box, a_tgv = make(box), make(tgv)
    steps = {}
    for a in (a_box, a_tgv):
        n = 0
        for _ in range(2):
            a.ExecuteTimeSynchronizationStep()
            n += a.last_rk_steps
        steps[a] = n
    b_box, b_tgv = make(box), make(tgv)
    batch = Batch()
    batch.add(b_box, steps[a_box])
    batch.add(b_tgv, steps[a_tgv])
    batch.step(max(steps.values()))
    batch.wait()
    one_box = run_batch_of_one(box, steps[a_box], damping=box.damping_primitive())
    one_tgv = run_batch_of_one(tgv, steps[a_tgv])
    assert np.array_equal(b_box.solution(), one_box.solution())
    assert np.array_equal(b_tgv.solution(), one_tgv.solution())
    assert rel(b_box.solution(), a_box.solution()) < 1e-14
    assert rel(b_tgv.solution(), a_tgv.solution()) < 1e-14


def test_failing_case_is_evicted(lib):
    from vcp3d_open_amd.solver import Batch, SolverClass, ACTIVE, FLAGS
    good_case, bad_case = TGVCase(n=4, p=4, curvilinear=True), TGVCase(n=3, p=4)
    print("[case] eviction: negative a^2 in one of two cases", flush=True)
    a_good, steps = run_alone(good_case, 2)
    good = SolverClass(good_case)
    good.ReadRestartSolution(good_case.initial_conservative())
    bad = SolverClass(bad_case)
    bad.ReadRestartSolution(bad_case.initial_conservative())
    u = bad.solution()
    u[5, 4, :] = 0.0
    u[5, 4, 0] = 1.0            # V4 > 0: no valid pressure, a^2 not a number in element 5
    bad.set_solution(u)
    batch = Batch()
    batch.add(good, steps)
    batch.add(bad, steps)
    batch.step(steps)
    st = batch.status()
    assert st[bad][ACTIVE] == 0 and st[bad][FLAGS] & 2, st[bad]
    assert np.array_equal(bad.solution(), u)
    assert np.array_equal(good.solution(), run_batch_of_one(good_case, steps).solution())
    assert rel(good.solution(), a_good.solution()) < 1e-14


def test_tiles_match_oracle_and_alone(lib, oracle_lib):
    from vcp3d_open_amd.solver import Batch, SolverClass, volume_tiles
    kws = [dict(n=3, p=2), dict(n=4, p=3, curvilinear=True), dict(n=4, p=4, curvilinear=True),
           dict(n=2, p=5), dict(n=4, p=6, curvilinear=True)]
    for kw in kws:
        case = TGVCase(**kw)
        K, M = case.p + 1, case.rule // 2 + 1
        for nt in volume_tiles(K, M)[1:]:
            print(f"[case] tile {nt} for K={K} M={M}", flush=True)
            a, steps = run_alone(case, 2, nt=nt)
            g = SolverClass(case)
            g.ReadRestartSolution(case.initial_conservative())
            batch = Batch(tiles={(K, M): nt})
            batch.add(g, steps)
            batch.step(steps)
            batch.wait()
            assert rel(g.solution(), a.solution()) < 1e-14
            o = oracle_lib.Oracle(case)
            o.set_restart(case.initial_conservative())
            for _ in range(2):
                o.sync_step()
            assert rel(g.solution(), o.modal()) < 1e-12, (kw, nt)


def test_case_joining_mid_run_matches_alone(lib):
    """A case added to a batch after the others have taken RK steps (as the
    serving loop adds arriving requests) ends bit-identical to the same case
    alone, and so do the cases that were already running."""
    from vcp3d_open_amd.solver import Batch, SolverClass
    first, late = TGVCase(n=3, p=3, mach=0.2), TGVCase(n=4, p=4, curvilinear=True)
    print("[case] a case joins a running batch after 5 RK steps", flush=True)

    def make(c):
        g = SolverClass(c)
        g.ReadRestartSolution(c.initial_conservative())
        return g

    a, b = make(first), make(late)
    batch = Batch()
    batch.add(a, 12)
    batch.step(5)
    batch.add(b, 9)
    batch.step(9)
    st = batch.status()
    assert st[a][STEPS_DONE_] == 12 and st[b][STEPS_DONE_] == 9
    assert np.array_equal(a.solution(), run_batch_of_one(first, 12).solution())
    assert np.array_equal(b.solution(), run_batch_of_one(late, 9).solution())
