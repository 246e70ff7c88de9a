"""The body force per unit volume on the GPU path ("Body force per unit
volume (x,y,z)", ElementClass.cpp:4492-4533), through the C ABI: the
curvilinear Taylor-Green vortex with a body force against the reference's
own outputs (tests/golden/tgv_bf_curv_n4_p*.npz), and an open box whose
sponge elements carry the damping and CML sources beside the body force
against the oracle. Tolerances as the other parity tests: residuals 1e-9,
states 1e-12 (relative L2)."""
import numpy as np
import pytest

from tests.golden.make_golden import load
from vcp3d_open_amd.cases import OpenBoxCase

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def lib():
    from vcp3d_open_amd.solver import load_library
    return load_library()


@pytest.mark.parametrize("name,steps", [("tgv_bf_curv_n4_p3", 3), ("tgv_bf_curv_n4_p4", 0)])
def test_body_force_against_reference(lib, name, steps):
    from vcp3d_open_amd.solver import SolverClass
    print(f"[case] body force golden {name}", flush=True)
    case, ref = load(name)
    g = SolverClass(case)
    g.ReadRestartSolution(case.initial_conservative())
    md = g.Residual(True)
    assert rel(g.residual_array(), ref["res0"]) < 1e-9
    for s in range(1, steps + 1):
        md = g.ExecuteTimeSynchronizationStep()
        if f"sol_{s}" in ref:
            assert rel(g.WriteRestartSolution(), ref[f"sol_{s}"]) < 1e-12
            assert rel(g.residual_array(), ref[f"res_{s}"]) < 1e-9
    assert abs(md[6] - ref["monitoring"][6]) <= 1e-12 * ref["monitoring"][6]
    g.close()


def test_body_force_with_sponge_against_oracle(lib, oracle_lib):
    """6 x 3 x 3 open box, P = 3, a 2-element sponge + CML at the outlet and a
    body force: the sponge elements project both sources together."""
    from vcp3d_open_amd.solver import SolverClass
    print("[case] body force + sponge open box", flush=True)
    case = OpenBoxCase(n=(6, 3, 3), p=3, sponge_elems=2, body_force=(2000.0, 500.0, -700.0))
    o = oracle_lib.Oracle(case)
    o.set_restart(case.initial_conservative())
    o.set_damping_state(case.damping_primitive())
    g = SolverClass(case)
    g.ReadRestartSolution(case.initial_conservative())
    g.SetDampingState(case.damping_primitive())
    o.residual()
    g.Residual(True)
    assert rel(g.residual_array(), o.residual_array()) < 1e-9
    for _ in range(2):
        o.sync_step()
        g.ExecuteTimeSynchronizationStep()
    assert rel(g.WriteRestartSolution(), o.nodal()) < 1e-12
    assert rel(g.residual_array(), o.residual_array()) < 1e-9
    g.close()
