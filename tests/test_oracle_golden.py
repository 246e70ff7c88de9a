"""The C oracle against the reference's own outputs (tests/golden/*.npz,
written by tests/golden/make_golden.py from the reference build).

Tolerances: fields after k RK steps within relative L2 1e-12; residuals
within 1e-9 (2e-8 for the 32^3 case). The residual R = M^-1 res of these
low-Mach flows is a small difference of large flux terms: a 1e-16 relative
perturbation of the modal input moves it by ~2e-12, and a 1e-15 relative
perturbation of the grid coordinates (the size of roundoff differences in
the metric terms of two correct FP64 set-ups) moves it by 1.4e-10 on 4^3 and
8.6e-9 on 32^3 elements (test_residual_conditioning measures both).
"""
import numpy as np
import pytest

from tests.golden.make_golden import load

SOL_TOL = 1e-12
RES_TOL = 1e-9
RES_TOL_32 = 2e-8


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def orc(oracle_lib):
    return oracle_lib


def run_oracle(orc, case, steps):
    o = orc.Oracle(case)
    o.set_restart(case.initial_conservative())
    o.residual()
    out = {"res0": o.residual_array()}
    for s in range(1, steps + 1):
        o.sync_step()
        out[f"sol_{s}"] = o.nodal()
        out[f"res_{s}"] = o.residual_array()
    return o, out


def test_reference_conditions(orc):
    case, g = load("tgv_curv_n4_p4")
    o = orc.Oracle(case)
    r = o.reference()
    np.testing.assert_allclose([r["u_ref"], r["p_ref"], r["rho_ref"], r["mu"]], g["reference"], rtol=1e-14)


@pytest.mark.parametrize("name,steps", [("tgv_cart_n3_p2", 0), ("tgv_curv_n4_p4", 1), ("tgv_curv_n4_p5", 4),
                                        ("tgv_curv_n4_p6", 0)])
def test_small_cases(orc, name, steps):
    case, g = load(name)
    _, out = run_oracle(orc, case, steps)
    assert rel(out["res0"], g["res0"]) < RES_TOL
    for key in g:
        if key.startswith("sol_"):
            assert rel(out[key], g[key]) < SOL_TOL, key
        if key.startswith("res_"):
            assert rel(out[key], g[key]) < RES_TOL, key


def test_config0_taylor_green_100_steps(orc):
    """8^3 Cartesian TGV, P = 3: residual at t = 0 and the state after 100
    SSP-RK3 steps."""
    case, g = load("tgv_cart_n8_p3")
    o = orc.Oracle(case)
    o.set_restart(case.initial_conservative())
    md = o.residual()
    assert rel(o.residual_array(), g["res0"]) < RES_TOL
    for _ in range(100):
        md = o.sync_step()
    assert rel(o.nodal(), g["sol_100"]) < SOL_TOL
    assert abs(md[6] - g["monitoring"][6]) < 1e-12 * abs(g["monitoring"][6])


@pytest.mark.slow
def test_config1_curvilinear_subset(orc):
    """32^3 curvilinear TGV, P = 4: every 331st element and the global
    per-variable norms of the residual and of the state after one step."""
    case, g = load("tgv_curv_n32_p4")
    sel = g["elements"]
    o = orc.Oracle(case)
    o.set_restart(case.initial_conservative())
    o.residual()
    r0 = o.residual_array()
    assert rel(r0[sel], g["res0"]) < RES_TOL_32
    np.testing.assert_allclose(np.sqrt((r0 ** 2).sum(axis=(0, 2))), g["res0_norms"], rtol=RES_TOL_32)
    o.sync_step()
    s1 = o.nodal()
    assert rel(s1[sel], g["sol_1"]) < SOL_TOL
    norms = np.sqrt((s1 ** 2).sum(axis=(0, 1)))
    scale = g["sol_1_norms"].copy()
    scale[1:4] = scale[1:4].max()          # momentum components share a scale
    assert np.all(np.abs(norms - g["sol_1_norms"]) < SOL_TOL * scale)
    assert rel(o.residual_array()[sel], g["res_1"]) < RES_TOL_32


def test_residual_conditioning(orc):
    """Sensitivity of R to a 1e-16 relative perturbation of the modal input
    and to a 1e-15 relative perturbation of the grid: bounds the agreement two
    correct FP64 implementations can reach."""
    case, _ = load("tgv_curv_n4_p4")
    o = orc.Oracle(case)
    o.set_restart(case.initial_conservative())
    U = o.modal()
    o.residual()
    R0 = o.residual_array()
    rng = np.random.default_rng(1)
    o.set_modal(U * (1.0 + 1e-16 * rng.standard_normal(U.shape)))
    o.residual()
    sens = rel(o.residual_array(), R0)
    assert 1e-14 < sens < 1e-10
    grid = case.element_grid() * (1.0 + 1e-15 * rng.standard_normal((case.n_elems, 3, case.n_dofs)))
    case.element_grid = lambda: grid
    o2 = orc.Oracle(case)
    o2.set_modal(U)
    o2.residual()
    sens_g = rel(o2.residual_array(), R0)
    assert 1e-12 < sens_g < 1e-9
