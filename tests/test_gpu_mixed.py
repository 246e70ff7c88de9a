"""Mixed-precision form (vcp_set_precision MIXED): FP32 sum-factorisation
contractions and mass-matrix solve, FP64 state, element means, pointwise
physics, metric terms, boundary kernels, time step and RK update.

Bars (measured values in the comments):
- the 8^3 P=4 curvilinear Taylor-Green trajectory over 100 RK steps stays
  within relative L2 1e-6 of the FP64 oracle over the whole nodal state
  (observed 7e-8), and its kinetic energy and enstrophy within 1e-6;
- the residual of the mixed form is within 2e-3 (relative L2) of the FP64
  residual at P = 2..6 and on an open box (observed 3e-5 .. 8.4e-4);
- in a batch, a mixed case is bit-identical in a batch of many and in a
  batch of one, and mixed and FP64 cases of the same shape form two groups.
"""
import numpy as np
import pytest

from vcp3d_open_amd.cases import OpenBoxCase, TGVCase

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_mixed_trajectory_against_fp64_oracle(oracle_lib):
    from vcp3d_open_amd.diagnostics import flow_integrals
    from vcp3d_open_amd.solver import SolverClass, host_setup
    case = TGVCase(n=8, p=4, curvilinear=True)
    print("[case] mixed 8^3 P=4 curvilinear TGV, 100 RK steps vs the FP64 oracle", flush=True)
    g = SolverClass(case, precision="mixed")
    g.ReadRestartSolution(case.initial_conservative())
    o = oracle_lib.Oracle(case)
    o.set_restart(case.initial_conservative())
    steps = 0
    while steps < 100:
        g.ExecuteTimeSynchronizationStep()
        o.sync_step()
        steps += g.last_rk_steps
    err = rel(g.WriteRestartSolution(), o.nodal())
    vol = host_setup(case)["vol"]
    fm, fo = flow_integrals(case, g.solution(), vol), flow_integrals(case, o.modal(), vol)
    e_ke = abs(fm["kinetic_energy"] / fo["kinetic_energy"] - 1)
    e_z = abs(fm["enstrophy"] / fo["enstrophy"] - 1)
    print(f"field rel L2 {err:.3e}, KE {e_ke:.3e}, enstrophy {e_z:.3e}", flush=True)
    assert err < 1e-6
    assert e_ke < 1e-6 and e_z < 1e-6


@pytest.mark.parametrize("kw", [dict(n=4, p=2), dict(n=4, p=3, curvilinear=True), dict(n=4, p=4, curvilinear=True),
                                dict(n=4, p=5, curvilinear=True), dict(n=4, p=6, curvilinear=True)],
                         ids=lambda kw: f"p{kw['p']}")
def test_mixed_residual_close_to_fp64(kw):
    from vcp3d_open_amd.solver import SolverClass
    case = TGVCase(**kw)
    print(f"[case] mixed vs fp64 residual {kw}", flush=True)
    r = {}
    for prec in ("fp64", "mixed"):
        g = SolverClass(case, precision=prec)
        g.ReadRestartSolution(case.initial_conservative())
        g.Residual(True)
        r[prec] = g.residual_array()
    assert rel(r["mixed"], r["fp64"]) < 2e-3


def test_mixed_open_box_residual():
    from vcp3d_open_amd.solver import SolverClass
    box = OpenBoxCase(n=(6, 2, 2), p=3, sponge_elems=2)
    print("[case] mixed vs fp64 residual, open box 6x2x2 P=3", flush=True)
    r = {}
    for prec in ("fp64", "mixed"):
        g = SolverClass(box, precision=prec)
        g.ReadRestartSolution(box.initial_conservative())
        g.SetDampingState(box.damping_primitive())
        g.Residual(True)
        r[prec] = g.residual_array()
    assert rel(r["mixed"], r["fp64"]) < 1e-3


def test_mixed_cases_in_a_batch():
    from vcp3d_open_amd.solver import Batch, SolverClass
    cases = [TGVCase(n=3, p=3, mach=0.2), TGVCase(n=4, p=4, curvilinear=True), TGVCase(n=2, p=5)]
    print("[case] batch of mixed and fp64 cases", flush=True)

    def make(c, prec):
        g = SolverClass(c, precision=prec)
        g.ReadRestartSolution(c.initial_conservative())
        return g

    batch = Batch()
    members = {}
    for c in cases:
        for prec in ("fp64", "mixed"):
            members[(id(c), prec)] = make(c, prec)
            batch.add(members[(id(c), prec)], 3)
    groups = batch.groups()
    assert len(groups) == 6 and sorted(x["mixed"] for x in groups) == [0, 0, 0, 1, 1, 1]
    batch.step(3)
    batch.wait()
    for c in cases:
        one = make(c, "mixed")
        b1 = Batch()
        b1.add(one, 3)
        b1.step(3)
        b1.wait()
        assert np.array_equal(members[(id(c), "mixed")].solution(), one.solution())
        assert rel(members[(id(c), "mixed")].solution(), members[(id(c), "fp64")].solution()) < 1e-7
