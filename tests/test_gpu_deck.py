"""The GPU path set up from case decks (vcp3d_open_amd/deck.py), through the
C ABI: the reference's own test decks (tests/golden/ref_n3_8x4x4/ and
ref_duct/, from VCP3D_CurviLinear/test/n3_8x4x4 and n5_15x8x8/duct) and the
wall variants of the first against the reference's outputs on them, and a
checkpoint written by write_restart resumed from a deck.

Tolerances as for the other parity tests: residual (R = M^-1 res) within
relative L2 1e-9, fields within 1e-12.
"""
import os

import numpy as np
import pytest

from tests.golden.make_golden import HERE as GOLDEN, load
from vcp3d_open_amd.deck import open_solver, write_restart

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def lib():
    from vcp3d_open_amd.solver import load_library
    return load_library()


@pytest.mark.parametrize("name,deck", [("deck_n3_8x4x4_static", "param_static.cfg"),
                                       ("deck_n3_8x4x4_total", "param_total.cfg"),
                                       ("deck_walls_a", "param_walls_a.cfg"),
                                       ("deck_walls_b", "param_walls_b.cfg"),
                                       ("deck_walls_c", "param_walls_c.cfg"),
                                       ("deck_walls_d", "param_walls_d.cfg"),
                                       ("deck_walls_e", "param_walls_e.cfg")])
def test_reference_deck_matches_reference(lib, name, deck):
    print(f"[case] {name}", flush=True)
    _, gold = load(name)
    s, case = open_solver(os.path.join(GOLDEN, "ref_n3_8x4x4", deck))
    np.testing.assert_allclose(list(s.reference().values()), gold["reference"], rtol=1e-14)
    s.Residual(True)
    assert rel(s.residual_array(), gold["res0"]) < 1e-9
    for _ in range(3):
        md = s.ExecuteTimeSynchronizationStep()
    # the reference driver prints the monitoring data of its last step: wall
    # force coefficients (closed walls cancel their pressure components from
    # terms of order 1e2 to ~1e-12, hence atol) and the maximum Mach^2
    np.testing.assert_allclose(md[:6], gold["monitoring"][:6], rtol=1e-8, atol=1e-10)
    assert abs(md[6] - gold["monitoring"][6]) <= 1e-9 * gold["monitoring"][6]
    assert rel(s.WriteRestartSolution(), gold["sol_3"]) < 1e-12
    assert rel(s.residual_array(), gold["res_3"]) < 1e-9


def test_reference_duct_deck(lib):
    """The reference's duct start deck: 15 x 8 x 8, P = 5, subsonic
    total-condition inflow, NSCBC outflow, four adiabatic walls; the state
    after 3 steps on every 13th element and the global conservative norms."""
    print("[case] deck_duct_start", flush=True)
    _, gold = load("deck_duct_start")
    s, case = open_solver(os.path.join(GOLDEN, "ref_duct", "extrapolation", "start.cfg"))
    s.Residual(True)
    for _ in range(3):
        md = s.ExecuteTimeSynchronizationStep()
    np.testing.assert_allclose(md[:6], gold["monitoring"][:6], rtol=1e-8, atol=1e-10)
    assert abs(md[6] - gold["monitoring"][6]) <= 1e-9 * gold["monitoring"][6]
    sol = s.WriteRestartSolution()
    assert rel(sol[gold["elements"]], gold["sol_3"][..., :5]) < 1e-12
    avg, n = s.averages()
    assert n == 3
    assert rel(avg[gold["elements"]], gold["sol_3"][..., 5:17]) < 1e-12
    # per-variable norms as one vector (the transverse momenta are ~1e-6 of it)
    assert rel(np.sqrt((sol ** 2).sum(axis=(0, 1))), gold["sol_3_norms"][:5]) < 1e-12


def test_checkpoint_and_resume(lib, tmp_path):
    """Two steps, a restart file, a new solver from a deck with "Restart: yes",
    one more step: the same state as three uninterrupted steps."""
    src = os.path.join(GOLDEN, "ref_n3_8x4x4")
    print("[case] deck checkpoint/resume", flush=True)
    a, case = open_solver(os.path.join(src, "param_static.cfg"))
    b, _ = open_solver(os.path.join(src, "param_static.cfg"))
    for _ in range(3):
        a.ExecuteTimeSynchronizationStep()
    for _ in range(2):
        b.ExecuteTimeSynchronizationStep()
    (tmp_path / "grid").mkdir()
    (tmp_path / "grid" / "Box_nPoly3_8x4x4.xyz").write_bytes(open(case.grid_file, "rb").read())
    (tmp_path / "init").mkdir()
    write_restart(b, case, str(tmp_path / "init" / "Restart.dat"), 2)
    text = open(os.path.join(src, "param_static.cfg")).read().replace("Restart: no", "Restart: yes")
    (tmp_path / "param.cfg").write_text(text)
    c, _ = open_solver(str(tmp_path / "param.cfg"))
    c.ExecuteTimeSynchronizationStep()
    assert rel(c.WriteRestartSolution(), a.WriteRestartSolution()) < 1e-12


def test_checkpoint_and_resume_with_averages(lib, tmp_path):
    """The duct deck (running averages on): two steps, a restart with the
    averaged columns, a new solver with "Restart: yes" and "Continue averaging
    from restart: yes", one more step: state and averages as after three
    uninterrupted steps."""
    src = os.path.join(GOLDEN, "ref_duct")
    print("[case] duct checkpoint/resume with averages", flush=True)
    deck = os.path.join(src, "extrapolation", "start.cfg")
    a, case = open_solver(deck)
    b, _ = open_solver(deck)
    md_a = [a.ExecuteTimeSynchronizationStep()[:6] for _ in range(3)]
    for _ in range(2):
        b.ExecuteTimeSynchronizationStep()
    # the running mean with weights 1/n of each step's force coefficients is
    # their arithmetic mean (the duct's four walls give nonzero coefficients)
    cf_a = a.force_averages()
    assert np.abs(cf_a).max() > 0.0
    np.testing.assert_allclose(cf_a, np.mean(md_a, axis=0), rtol=1e-13, atol=1e-14 * np.abs(md_a).max())
    (tmp_path / "grid").mkdir()
    (tmp_path / "grid" / "Box_nPoly5_15x8x8.xyz").write_bytes(open(case.grid_file, "rb").read())
    run = tmp_path / "extrapolation"
    (run / "init").mkdir(parents=True)
    write_restart(b, case, str(run / "init" / "Restart.dat"), 2)
    # the restart's meta data carry b's averaged force coefficients
    from vcp3d_open_amd import io as vio
    *_, meta = vio.read_su2_restart(str(run / "init" / "Restart.dat"), case.n_dofs, with_meta=True)
    assert np.array_equal(meta[:6], b.force_averages()) and not meta[6:].any()
    text = open(deck).read().replace("Restart: no", "Restart: yes")
    text = text.replace("Continue averaging from restart: no", "Continue averaging from restart: yes")
    (run / "start.cfg").write_text(text)
    c, c_case = open_solver(str(run / "start.cfg"))
    assert c_case.continue_averaging
    assert np.array_equal(c.force_averages(), meta[:6])
    c.ExecuteTimeSynchronizationStep()
    # continued, not restarted from zero: the mean over all three steps (the
    # resumed state agrees to 1e-12; the walls' pressure components are
    # cancellation residues of ~1e-13 from terms of order 1e2, hence atol)
    np.testing.assert_allclose(c.force_averages(), cf_a, rtol=1e-9, atol=1e-12)
    assert rel(c.WriteRestartSolution(), a.WriteRestartSolution()) < 1e-12
    avg_a, n_a = a.averages()
    avg_c, n_c = c.averages()
    assert n_a == n_c == 3
    assert rel(avg_c, avg_a) < 1e-12
