"""Launch timing inside the synchronization step (vcp_set_kernel_timing /
vcp_get_kernel_timing, through the C ABI): the timed steps give the same
state as untimed ones, every RK stage is counted once, and the face and
volume durations are positive and fit inside the wall time of the steps."""
import time

import numpy as np
import pytest

from vcp3d_open_amd.cases import TGVCase

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    from vcp3d_open_amd.solver import load_library
    return load_library()


def test_timed_steps_match_untimed_and_count_stages(lib):
    from vcp3d_open_amd.solver import SolverClass
    print("[case] kernel timing tgv_curv_n4_p4", flush=True)
    case = TGVCase(n=4, p=4, curvilinear=True)
    runs = []
    for timed in (False, True):
        s = SolverClass(case)
        s.ReadRestartSolution(case.initial_conservative())
        s.Residual(True)
        s.set_kernel_timing(timed)
        rk = 0
        t0 = time.perf_counter()
        for _ in range(3):
            s.ExecuteTimeSynchronizationStep()
            rk += s.last_rk_steps
        wall_ms = (time.perf_counter() - t0) * 1e3
        face, vol, stages = s.kernel_timing()
        runs.append((s.WriteRestartSolution(), rk, face, vol, stages, wall_ms))
        s.close()
    (a, _, _, _, n_off, _), (b, rk, face, vol, n_on, wall_ms) = runs
    assert np.array_equal(a, b)
    assert n_off == 0
    assert n_on == 3 * rk
    assert 0.0 < face and 0.0 < vol
    assert n_on * (face + vol) < wall_ms


def test_timing_stops_and_keeps_its_sums(lib):
    from vcp3d_open_amd.solver import SolverClass
    print("[case] kernel timing stop/restart", flush=True)
    case = TGVCase(n=3, p=2)
    s = SolverClass(case)
    s.ReadRestartSolution(case.initial_conservative())
    s.set_kernel_timing(True)
    s.ExecuteTimeSynchronizationStep()
    first = s.kernel_timing()
    s.set_kernel_timing(False)
    s.ExecuteTimeSynchronizationStep()
    assert s.kernel_timing() == first          # stopped: no further stages
    s.set_kernel_timing(True)                  # restarted: sums cleared
    assert s.kernel_timing()[2] == 0
    s.ExecuteTimeSynchronizationStep()
    assert s.kernel_timing()[2] == 3 * s.last_rk_steps
    s.close()
