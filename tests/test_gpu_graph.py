"""The RK step replayed from captured HIP graphs (vcp_set_graph) against the
launch-by-launch path, through the C ABI.

Both paths run the same kernels on the same inputs in the same order, so the
bar is bit-identical states, residuals and monitoring data after every
synchronization step. The two solvers alternate, so every graph replay is
separated from the next by eager launches and host synchronizations of the
other context, and every replay is checked. The launch-by-launch path itself
is checked against the oracle in test_gpu_parity.py and test_gpu_openbox.py;
one oracle comparison here ties the graph path to it directly.
"""
import numpy as np
import pytest

from vcp3d_open_amd.cases import OpenBoxCase, TGVCase

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    from vcp3d_open_amd.solver import load_library
    return load_library()


def pair(case):
    from vcp3d_open_amd.solver import SolverClass
    eager, graph = SolverClass(case), SolverClass(case, use_graph=True)
    nodal = case.initial_conservative()
    for g in (eager, graph):
        g.ReadRestartSolution(nodal)
        if getattr(case, "sponge_elems", 0) > 0:
            g.SetDampingState(case.damping_primitive())
    return eager, graph


def step_and_compare(eager, graph, steps):
    for s in range(steps):
        md_e = eager.ExecuteTimeSynchronizationStep()
        md_g = graph.ExecuteTimeSynchronizationStep()
        assert np.array_equal(graph.solution(), eager.solution()), f"state differs after step {s + 1}"
        assert np.array_equal(graph.residual_array(), eager.residual_array()), f"residual differs after step {s + 1}"
        assert np.array_equal(md_g, md_e)


def test_graph_matches_eager_periodic(lib, oracle_lib):
    case = TGVCase(n=4, p=4, curvilinear=True)
    print("[case] graph curv_n4_p4 6 steps", flush=True)
    eager, graph = pair(case)
    step_and_compare(eager, graph, 6)
    o = oracle_lib.Oracle(case)
    o.set_restart(case.initial_conservative())
    for _ in range(6):
        o.sync_step()
    rel = np.linalg.norm(graph.solution() - o.modal()) / np.linalg.norm(o.modal())
    assert rel < 1e-12, rel


def test_graph_matches_eager_several_rk_steps_per_call(lib):
    """Synchronization step of about 2.5 stable time steps: the captures of the
    inner steps (not the last of a call) are replayed too, from both state
    buffers."""
    kw = dict(n=3, p=3)
    probe, _ = pair(TGVCase(**kw))
    dt_min = probe.time_steps().min() / probe.reference()["u_ref"]
    case = TGVCase(**kw, dt_synchr=2.5 * dt_min)
    print("[case] graph cart_n3_p3 multi-step synchronization", flush=True)
    eager, graph = pair(case)
    step_and_compare(eager, graph, 4)
    # a new state between calls: the next call recomputes its time step
    nodal = case.initial_conservative()
    eager.ReadRestartSolution(nodal)
    graph.ReadRestartSolution(nodal)
    step_and_compare(eager, graph, 1)


def test_graph_matches_eager_open_box(lib):
    """PC-NSCBC faces, the Mach averaging and the sponge + CML sources inside
    the captured step."""
    case = OpenBoxCase(n=(6, 2, 2), p=3, sponge_elems=2)
    print("[case] graph open box 6x2x2 p3", flush=True)
    eager, graph = pair(case)
    step_and_compare(eager, graph, 4)
