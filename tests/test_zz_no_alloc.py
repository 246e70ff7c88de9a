"""vcp_residual_static through the C ABI: the residual evaluated with the
face-slot buffer in the library's fixed device array (no allocation in the
call) equals vcp_residual's, bit for bit, on two calls in a row (the second
on a smaller case, so stale contents of the array would show), and a case
whose buffer exceeds the array is rejected with an error."""
import ctypes as C

import numpy as np
import pytest

from vcp3d_open_amd.cases import OpenBoxCase, TGVCase

pytestmark = pytest.mark.gpu


def residual_static(solver, monitor=True):
    from vcp3d_open_amd.solver import load_library
    L = load_library()
    md = np.zeros(8)
    rc = L.vcp_residual_static(solver.h, 1 if monitor else 0, md)
    return rc, md


def run_pair(case, damping=None):
    from vcp3d_open_amd.solver import SolverClass
    a, b = SolverClass(case), SolverClass(case)
    for g in (a, b):
        g.ReadRestartSolution(case.initial_conservative())
        if damping is not None:
            g.SetDampingState(damping)
    md_a = a.Residual(True)
    rc, md_b = residual_static(b)
    assert rc == 0, b._L.vcp_last_error().decode()
    return a, b, md_a, md_b


def test_static_residual_matches_residual():
    cases = [("curvilinear TGV 4^3 P=4", TGVCase(n=4, p=4, curvilinear=True), None),
             ("TGV 3^3 P=3", TGVCase(n=3, p=3), None)]
    box = OpenBoxCase(n=(6, 2, 2), p=3, sponge_elems=2)
    cases.append(("open box 6x2x2 P=3", box, box.damping_primitive()))
    for label, case, damping in cases:
        print(f"[case] vcp_residual_static vs vcp_residual: {label}", flush=True)
        a, b, md_a, md_b = run_pair(case, damping)
        assert np.array_equal(b.residual_array(), a.residual_array()), label
        assert np.array_equal(md_b, md_a), label
        # the state is untouched and the context goes on as after vcp_residual
        assert np.array_equal(b.solution(), a.solution())
        a.ExecuteTimeSynchronizationStep()
        b.ExecuteTimeSynchronizationStep()
        assert np.array_equal(b.solution(), a.solution()), label


def test_static_residual_rejects_oversized_case():
    from vcp3d_open_amd.solver import SolverClass
    case = TGVCase(n=18, p=4)          # 5832 elements: 70 MB of face slots > 64 MiB
    g = SolverClass(case)
    g.ReadRestartSolution(case.initial_conservative())
    print("[case] vcp_residual_static on 18^3 P=4 (rejected before any launch)", flush=True)
    rc, _ = residual_static(g)
    assert rc == -7
    assert "exceeds the fixed scratch array" in g._L.vcp_last_error().decode()
