"""Subgrid scale models (WALE, Vreman) on the GPU path, through the C ABI:
against the oracle on the same inputs and against the reference's own
outputs on cases with "Subgrid scale model: WALE / Vreman".

Tolerances as for the other parity tests: the residual within relative L2
3e-11 when the oracle runs on the GPU path's own metric arrays (1e-9 against
the reference's files), fields after k RK steps within 1e-12, element time
steps within 1e-13; the monitored maximum eddy viscosity (monitoring[7])
within 1e-10 of the oracle and 1e-9 of the reference. The duct deck's
eddy-viscosity average and maximum are compared within 1e-6: its WALE eddy
viscosity is set by transverse velocities ~1e-6 of the streamwise one (the
oracle's test of the same case gives the reason and the observed 4e-8).
"""
import os

import numpy as np
import pytest

from tests.golden.make_golden import HERE as GOLDEN, load
from vcp3d_open_amd.cases import TGVCase

pytestmark = pytest.mark.gpu

CASES = [
    ("wale_curv_n4_p3", dict(n=4, p=3, curvilinear=True, sgs="wale")),
    ("vreman_curv_n4_p4", dict(n=4, p=4, curvilinear=True, sgs="vreman")),
    ("wale_curv_n4_p4", dict(n=4, p=4, curvilinear=True, sgs="wale")),
    ("wale_cart_n3_p2", dict(n=3, p=2, sgs="wale")),
    ("vreman_curv_n4_p5", dict(n=4, p=5, curvilinear=True, sgs="vreman")),
    ("wale_curv_n4_p6", dict(n=4, p=6, curvilinear=True, sgs="wale")),
]


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def lib():
    from vcp3d_open_amd.solver import load_library
    return load_library()


@pytest.mark.parametrize("name,kw", CASES, ids=[c[0] for c in CASES])
def test_sgs_residual_on_shared_geometry(lib, oracle_lib, name, kw):
    from vcp3d_open_amd.solver import SolverClass, host_setup
    case = TGVCase(**kw)
    print(f"[case] sgs {name} residual", flush=True)
    o = oracle_lib.Oracle(case)
    o.use_geometry(host_setup(case))
    g = SolverClass(case)
    nodal = case.initial_conservative()
    o.set_restart(nodal)
    g.ReadRestartSolution(nodal)
    md_o = o.residual()
    md_g = g.Residual(True)
    err = rel(g.residual_array(), o.residual_array())
    print(f"[case] sgs {name}: residual rel {err:.3e}, eddy max {md_g[7]:.6e} (oracle {md_o[7]:.6e})", flush=True)
    assert err < 3e-11
    assert md_o[7] > 0.0
    assert abs(md_g[7] - md_o[7]) <= 1e-10 * md_o[7]


@pytest.mark.parametrize("name,kw", [CASES[0], CASES[1], CASES[4]], ids=[CASES[i][0] for i in (0, 1, 4)])
def test_sgs_time_steps_and_rk_steps(lib, oracle_lib, name, kw):
    """Element time steps with the nodal eddy viscosity, then three
    synchronization steps."""
    from vcp3d_open_amd.solver import SolverClass
    case = TGVCase(**kw)
    print(f"[case] sgs {name} steps", flush=True)
    o = oracle_lib.Oracle(case)
    g = SolverClass(case)
    nodal = case.initial_conservative()
    o.set_restart(nodal)
    g.ReadRestartSolution(nodal)
    dto, dtg = o.time_steps(), g.time_steps()
    assert rel(dtg, dto) < 1e-13
    # the model lowers the time step against the same case without it
    plain = TGVCase(**dict(kw, sgs="none"))
    p = oracle_lib.Oracle(plain)
    p.set_restart(nodal)
    assert np.all(dto <= p.time_steps()) and np.any(dto < p.time_steps())
    for _ in range(3):
        o.sync_step()
        g.ExecuteTimeSynchronizationStep()
    assert rel(g.solution(), o.modal()) < 1e-12
    assert rel(g.residual_array(), o.residual_array()) < 1e-9


@pytest.mark.parametrize("name,steps", [("tgv_wale_curv_n4_p3", 3), ("tgv_vreman_curv_n4_p5", 2),
                                        ("tgv_wale_curv_n4_p4", 0), ("tgv_vreman_curv_n4_p4", 0)])
def test_sgs_tgv_against_reference(lib, name, steps):
    from vcp3d_open_amd.solver import SolverClass
    print(f"[case] sgs golden {name}", flush=True)
    case, ref = load(name)
    g = SolverClass(case)
    g.ReadRestartSolution(case.initial_conservative())
    md = g.Residual(True)
    assert rel(g.residual_array(), ref["res0"]) < 1e-9
    for s in range(1, steps + 1):
        md = g.ExecuteTimeSynchronizationStep()
        if f"sol_{s}" in ref:
            assert rel(g.WriteRestartSolution(), ref[f"sol_{s}"]) < 1e-12
        if f"res_{s}" in ref:
            assert rel(g.residual_array(), ref[f"res_{s}"]) < 1e-9
    assert abs(md[7] - ref["monitoring"][7]) <= 1e-9 * ref["monitoring"][7]


@pytest.mark.parametrize("name,deck", [("deck_walls_a_wale", "param_walls_a_wale.cfg"),
                                       ("deck_walls_b_vreman", "param_walls_b_vreman.cfg")])
def test_sgs_wall_decks_against_reference(lib, name, deck):
    """The 8 x 4 x 4 wall decks with an SGS model: isothermal walls with
    subsonic inflow/outflow (WALE), NSCBC inflow/outflow with an inviscid and
    a heat-flux wall (Vreman); the eddy viscosity enters the wall and NSCBC
    boundary faces with the element's own LES length scale."""
    from vcp3d_open_amd.deck import open_solver
    print(f"[case] sgs deck {name}", flush=True)
    _, gold = load(name)
    s, _ = open_solver(os.path.join(GOLDEN, "ref_n3_8x4x4", deck))
    s.Residual(True)
    assert rel(s.residual_array(), gold["res0"]) < 1e-9
    for _ in range(3):
        md = s.ExecuteTimeSynchronizationStep()
    assert rel(s.WriteRestartSolution(), gold["sol_3"]) < 1e-12
    assert rel(s.residual_array(), gold["res_3"]) < 1e-9
    assert abs(md[6] - gold["monitoring"][6]) <= 1e-9 * gold["monitoring"][6]
    assert abs(md[7] - gold["monitoring"][7]) <= 1e-9 * gold["monitoring"][7]


def test_sgs_duct_deck_averages(lib):
    """The duct deck with WALE: the state, the running averages and the
    averaged eddy viscosity against the reference's averaged restart columns."""
    from vcp3d_open_amd.deck import open_solver
    print("[case] sgs deck_duct_wale", flush=True)
    _, gold = load("deck_duct_wale")
    sel = gold["elements"]
    s, _ = open_solver(os.path.join(GOLDEN, "ref_duct", "extrapolation", "start_wale.cfg"))
    s.Residual(True)
    for _ in range(3):
        md = s.ExecuteTimeSynchronizationStep()
    assert rel(s.WriteRestartSolution()[sel], gold["sol_3"][..., :5]) < 1e-12
    avg, n = s.averages()
    assert n == 3
    assert rel(avg[sel][..., :11], gold["sol_3"][..., 5:16]) < 1e-12
    assert rel(avg[sel][..., 11], gold["sol_3"][..., 16]) < 1e-6
    assert abs(md[7] - gold["monitoring"][7]) <= 1e-6 * gold["monitoring"][7]


def one_step_per_sync(oracle_lib, kw):
    """The case with its synchronization time step 0.7 times its smallest
    element time step: every synchronization step is one RK step."""
    probe = oracle_lib.Oracle(TGVCase(**kw))
    probe.set_restart(TGVCase(**kw).initial_conservative())
    dt_min = probe.time_steps().min() / probe.reference()["u_ref"]
    return TGVCase(**kw, dt_synchr=0.7 * dt_min)


def test_sgs_batch_of_one_and_mixed(lib, oracle_lib):
    """A WALE case and a Vreman case in one batch with a case without a model
    (three groups): each equals its batch of one bit for bit and the oracle to
    1e-12; the mixed-precision form of the WALE case stays within 1e-6 of
    FP64 after the same steps."""
    from vcp3d_open_amd.solver import Batch, SolverClass, SYNC_DONE
    kws = [dict(n=4, p=3, curvilinear=True, sgs="wale"), dict(n=4, p=3, curvilinear=True, sgs="vreman"),
           dict(n=4, p=3, curvilinear=True)]
    cases = [one_step_per_sync(oracle_lib, kw) for kw in kws]
    steps = 4
    print("[case] sgs batch of three", flush=True)

    def batch_run(case_list, precision="fp64"):
        b = Batch()
        gs = []
        for c in case_list:
            g = SolverClass(c, precision=precision)
            g.ReadRestartSolution(c.initial_conservative())
            b.add(g, steps)
            gs.append(g)
        b.step(steps)
        st = b.status()
        b.close()
        return gs, st

    members, st = batch_run(cases)
    for c, g in zip(cases, members):
        assert st[g][SYNC_DONE] == steps, c.sgs
        (one,), _ = batch_run([c])
        assert np.array_equal(g.solution(), one.solution()), c.sgs
        o = oracle_lib.Oracle(c)
        o.set_restart(c.initial_conservative())
        for _ in range(steps):
            o.sync_step()
        assert rel(g.solution(), o.modal()) < 1e-12, c.sgs
    (mixed,), _ = batch_run([cases[0]], precision="mixed")
    err = rel(mixed.solution(), members[0].solution())
    print(f"[case] sgs mixed precision rel {err:.3e}", flush=True)
    assert err < 1e-6
