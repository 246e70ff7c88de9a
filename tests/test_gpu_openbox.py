"""GPU parity of the open-boundary terms: PC-NSCBC inflow/outflow, the
boundary Mach averaging, the sponge layer and the characteristic matching
layer, through the C ABI, against the C oracle on the same inputs and
against the reference's own outputs (tests/golden/ob_*.npz).

Tolerances as for the periodic path: residual (R = M^-1 res) within
relative L2 1e-9, fields after k RK steps within 1e-12, boundary-averaged
Mach numbers within 1e-13.
"""
import numpy as np
import pytest

from tests.golden.make_golden import load
from vcp3d_open_amd.cases import GAMMA, OpenBoxCase

pytestmark = pytest.mark.gpu

SMALL = ["ob_static_p3", "ob_total_p3", "ob_total_p2", "ob_static_p4", "ob_static_p5", "ob_static_p6"]


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def lib():
    from vcp3d_open_amd.solver import load_library
    return load_library()


def start_gpu(case, bug=True):
    from vcp3d_open_amd.solver import SolverClass
    g = SolverClass(case, bug_compat=bug)
    g.ReadRestartSolution(case.initial_conservative())
    if case.sponge_elems > 0:
        g.SetDampingState(case.damping_primitive())
    return g


def start_oracle(orc, case, bug=True):
    o = orc.Oracle(case, bug_compat=bug)
    o.set_restart(case.initial_conservative())
    if case.sponge_elems > 0:
        o.set_damping_state(case.damping_primitive())
    return o


@pytest.mark.parametrize("name", SMALL)
@pytest.mark.parametrize("bug", [True, False], ids=["bugcompat", "fixed"])
def test_open_box_matches_oracle(lib, oracle_lib, name, bug):
    case, _ = load(name)
    print(f"[case] {name} bug={bug}", flush=True)
    o, g = start_oracle(oracle_lib, case, bug), start_gpu(case, bug)
    np.testing.assert_allclose(list(g.reference().values()), list(o.reference().values()), rtol=1e-15)
    np.testing.assert_allclose(g.AverageDataNSCBC(), o.nscbc_average(), rtol=1e-13, atol=1e-16)
    md_o, md_g = o.residual(), g.Residual(True)
    ro, rg = o.residual_array(), g.residual_array()
    assert rel(rg, ro) < 1e-9, rel(rg, ro)
    i = np.arange(case.n_elems) % case.n[0]
    for sel in (i == 0, i == case.n[0] - 1):
        assert rel(rg[sel], ro[sel]) < 1e-9
    assert abs(md_g[6] - md_o[6]) <= 1e-12 * abs(md_o[6])
    for _ in range(2):
        o.sync_step()
        g.ExecuteTimeSynchronizationStep()
    assert rel(g.WriteRestartSolution(), o.nodal()) < 1e-12
    assert rel(g.residual_array(), o.residual_array()) < 1e-9


@pytest.mark.parametrize("name", SMALL)
def test_open_box_matches_reference(lib, name):
    case, gold = load(name)
    print(f"[case] {name} golden", flush=True)
    g = start_gpu(case)
    g.Residual(True)
    assert rel(g.residual_array(), gold["res0"]) < 1e-9
    steps = max([int(k.split("_")[1]) for k in gold if k.startswith("sol_")] or [0])
    for _ in range(steps):
        g.ExecuteTimeSynchronizationStep()
    if steps:
        assert rel(g.WriteRestartSolution(), gold[f"sol_{steps}"]) < 1e-12
        assert rel(g.residual_array(), gold[f"res_{steps}"]) < 1e-9


def test_open_box_full_size_against_reference(lib):
    """48 x 24 x 24, P = 4, M = 0.5, 8-element sponge + CML at the outlet:
    every 97th element and the per-variable norms of R(U_0)."""
    case, gold = load("ob_n48_p4")
    print("[case] ob_n48_p4 golden", flush=True)
    g = start_gpu(case)
    g.Residual(True)
    R = g.residual_array()
    assert rel(R[gold["elements"]], gold["res0"]) < 1e-9
    np.testing.assert_allclose(np.sqrt((R ** 2).sum(axis=(0, 2))), gold["res0_norms"], rtol=1e-9)


@pytest.mark.parametrize("inflow", ["static", "total"])
def test_open_box_free_stream(lib, inflow):
    """The matching free stream is a steady state of the whole open box."""
    case = OpenBoxCase(n=(6, 2, 2), p=3, inflow=inflow, sponge_elems=2)
    print(f"[case] free stream {inflow}", flush=True)
    q = case.damping_primitive()
    cons = np.stack([q[..., 0], q[..., 0] * q[..., 1], 0 * q[..., 0], 0 * q[..., 0],
                     q[..., 4] / (GAMMA - 1) + 0.5 * q[..., 0] * q[..., 1] ** 2], axis=-1)
    from vcp3d_open_amd.solver import SolverClass
    g = SolverClass(case)
    g.ReadRestartSolution(cons)
    g.SetDampingState(q)
    g.Residual(True)
    assert np.abs(g.residual_array()).max() < 1e-10
    np.testing.assert_allclose(g.AverageDataNSCBC(), [case.mach, case.mach, 0, 0, 0, 0], rtol=1e-12, atol=0)


def reflected_amplitude(case, nodal, x_range=(1.3, 5.0 / 3.0)):
    """Largest |p - p_inf| / (pulse amplitude p_inf) at the nodes between x0
    and the start of the outlet sponge: after the pulse's downstream half has
    left through the NSCBC outlet, what remains there is reflected."""
    x = case.nodal_coords()[..., 0]
    rho, ru, rv, rw, rE = (nodal[..., i] for i in range(5))
    p = (GAMMA - 1.0) * (rE - 0.5 * (ru ** 2 + rv ** 2 + rw ** 2) / rho)
    sel = (x >= x_range[0]) & (x < x_range[1])
    return float(np.abs(p[sel] - case.p_inf).max() / (case.pulse_amp * case.p_inf))


def test_pulse_reflection_against_reference(lib):
    """A 1e-3 pressure pulse in a 12 x 2 x 2, P = 3, M = 0.5 open box
    (PC-NSCBC static inflow/outflow, 2-element sponge + CML): after 500 and
    1000 RK steps the GPU state equals the reference's within relative L2
    1e-11, and the reflected-wave amplitude in front of the outlet sponge
    at step 1000 (about 5 % of the pulse) equals the reference's to 1e-8."""
    case, fx = load("ob_reflect_p3")
    print("[case] ob_reflect_p3: 1000 RK steps against the reference", flush=True)
    g = start_gpu(case)
    steps = 0
    for target in (500, 1000):
        while steps < target:
            g.ExecuteTimeSynchronizationStep()
            steps += g.last_rk_steps
        assert steps == target
        got, ref = g.WriteRestartSolution(), fx[f"sol_{target}"]
        err = rel(got, ref)
        print(f"step {target}: state rel {err:.3e}", flush=True)
        assert err < 1e-11
    a_g, a_r = reflected_amplitude(case, got), reflected_amplitude(case, ref)
    print(f"reflected amplitude {a_g:.6e} (reference {a_r:.6e})", flush=True)
    assert 0.01 < a_r < 0.2
    assert abs(a_g / a_r - 1) < 1e-8


def test_sponge_target_from_running_averages(lib):
    """The sponge target taken on the device from the case's own running
    averages (vcp_damping_from_averages) equals the host conversion of the
    same averaged primitives handed over as the averaged file's columns
    (vcp_set_damping_state, the reference's ReadAveragedSolutionSpongeLayer
    + InitializeDampingState path): the residuals agree to 1e-12, and differ
    from the residual with the original free-stream target."""
    case = OpenBoxCase(n=(6, 2, 2), p=3, sponge_elems=2, compute_average=True, pulse_amp=0.05, pulse_width=0.2)
    print("[case] open box 6x2x2 P=3, sponge target from 3 averaged steps", flush=True)
    a = start_gpu(case)
    for _ in range(3):
        a.ExecuteTimeSynchronizationStep()
    avg, n = a.averages()
    assert n == 3
    a.Residual(True)
    r_free = a.residual_array()
    a.DampingStateFromAverages()
    a.Residual(True)
    r_dev = a.residual_array()
    b = start_gpu(case)
    b.set_solution(a.solution())
    b.SetDampingState(np.ascontiguousarray(avg[..., :5]))
    b.Residual(True)
    r_host = b.residual_array()
    assert rel(r_dev, r_host) < 1e-12, rel(r_dev, r_host)
    assert rel(r_dev, r_free) > 1e-9
