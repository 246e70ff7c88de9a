"""Host-side pieces of the library and the package, no GPU: the set-up the
GPU path runs on (tables, metrics, neighbour maps) against the oracle, and
the file formats."""
import os

import numpy as np
import pytest

from vcp3d_open_amd import io as vio
from vcp3d_open_amd.cases import TGVCase
from vcp3d_open_amd.solver import host_setup


# n = 8 (512 elements) runs the element loops on several host threads
@pytest.mark.parametrize("kw", [dict(n=3, p=2), dict(n=4, p=4, curvilinear=True), dict(n=4, p=5, curvilinear=True),
                                dict(n=8, p=3, curvilinear=True)])
def test_host_setup_matches_oracle(oracle_lib, kw):
    case = TGVCase(**kw)
    h = host_setup(case)
    o = oracle_lib.Oracle(case)
    vol = o.setup(0)
    scale = np.abs(vol).max(axis=(0, 2), keepdims=True)
    assert np.all(np.abs(h["vol"] - vol) <= 1e-12 * scale)
    for f in range(6):
        fm = o.setup(1 + f)
        scale = np.abs(fm).max(axis=(0, 2), keepdims=True) + 1e-300
        assert np.all(np.abs(h["faces"][f] - fm) <= 1e-12 * scale), f
    np.testing.assert_allclose(h["lens"], o.setup(7), rtol=1e-12)
    assert np.array_equal(h["nbr"], o.setup(8).astype(np.int32))     # bit-exact maps


def test_batched_test_cases_pass_the_grid_checks():
    """The seeded cases of the batched GPU tests are accepted by the grid
    reader's checks (perturbed coarse grids fail its LGL spacing test)."""
    from tests.test_gpu_batch import seeded_cases
    for seed, count in ((7, 7), (11, 6)):
        for kw in seeded_cases(seed=seed, count=count):
            host_setup(TGVCase(**kw))


def test_unsupported_grid_is_rejected():
    from vcp3d_open_amd.solver import VCPError
    case = TGVCase(n=3, p=5, curvilinear=True)   # perturbation too strong for LGL detection
    with pytest.raises(VCPError, match="point distribution"):
        host_setup(case)


def test_plot3d_round_trip(tmp_path):
    case = TGVCase(n=3, p=3, curvilinear=True)
    xyz = case.grid_block()
    path = os.path.join(tmp_path, "g.xyz")
    vio.write_plot3d(path, xyz)
    assert np.array_equal(vio.read_plot3d(path), xyz)
    # header layout of the reference reader: 4 | 1 | 4 | 12 | il jl kl | 12 | 8*3*n
    raw = np.fromfile(path, dtype=np.int32, count=9)
    il = 3 * 3 + 1
    assert list(raw) == [4, 1, 4, 12, il, il, il, 12, 8 * 3 * il ** 3]


def test_su2_restart_round_trip(tmp_path):
    case = TGVCase(n=2, p=3)
    cons = case.initial_conservative()
    path = os.path.join(tmp_path, "r.dat")
    vio.write_su2_restart(path, case.nodal_coords(), cons, time_step=7)
    names, data, step = vio.read_su2_restart(path, case.n_dofs)
    assert names[:4] == ["x", "y", "z", "Density"] and step == 7
    assert np.array_equal(data[:, :, 3:], cons)
    hdr = np.fromfile(path, dtype=np.int32, count=5)
    assert list(hdr) == [535532, 8, case.n_elems * case.n_dofs, 1, 8]
    # meta data: zeros by default, the averaged force coefficients when given
    assert not vio.read_su2_restart(path, case.n_dofs, with_meta=True)[3].any()
    cf = np.array([0.1, -0.2, 3e-12, 0.04, 5e-3, -6e-4])
    vio.write_su2_restart(path, case.nodal_coords(), cons, time_step=7, meta=cf)
    *_, step, meta = vio.read_su2_restart(path, case.n_dofs, with_meta=True)
    assert step == 7 and np.array_equal(meta[:6], cf) and not meta[6:].any()
    # the meta data are the last 64 bytes of the file, after the step
    assert np.array_equal(np.frombuffer(open(path, "rb").read()[-64:], dtype="<f8"), np.r_[cf, 0.0, 0.0])


def test_taylor_green_initial_state():
    case = TGVCase(n=4, p=3)
    x = case.nodal_coords()
    cons = case.initial_conservative()
    rho = cons[..., 0]
    u = cons[..., 1] / rho
    v0 = case.ref["v_inf"]
    np.testing.assert_allclose(u, v0 * np.sin(x[..., 0]) * np.cos(x[..., 1]) * np.cos(x[..., 2]), atol=1e-12 * v0)
    assert np.abs(cons[..., 3]).max() == 0.0
    # periodic grid: opposite faces coincide after the translation
    g = case.grid_block()
    np.testing.assert_allclose(g[0, :, :, -1] - g[0, :, :, 0], 2 * np.pi, rtol=1e-15)


def test_open_box_host_setup_and_rejections():
    """Open box: physical faces have no neighbour; invalid boundary and
    sponge settings are refused by vcp_create before any device call."""
    import ctypes as C
    from vcp3d_open_amd.cases import OpenBoxCase
    from vcp3d_open_amd.solver import desc_from_case, host_setup, load_library
    case = OpenBoxCase(n=(4, 2, 3), p=3, sponge_elems=1)
    nbr = host_setup(case)["nbr"]
    i = np.arange(case.n_elems) % 4
    assert np.all(nbr[i == 0, 0] == -1) and np.all(nbr[i == 3, 1] == -1)
    assert np.all(nbr[:, 2:] >= 0) and np.all(nbr[i > 0, 0] >= 0)
    L = load_library()
    grid = np.ascontiguousarray(case.element_grid())
    for field, value, msg in [("outlet_average_local", 1, "averaged Mach"),
                              ("sponge_layer", 0, "sponge"),
                              ("periodic", (1, 1, 1), "periodic direction")]:
        d = desc_from_case(case)
        if isinstance(value, tuple):
            getattr(d, field)[:] = list(value)
        else:
            setattr(d, field, value)
        h = C.c_void_p()
        assert L.vcp_create(C.byref(d), grid, C.byref(h)) < 0
        assert msg in L.vcp_last_error().decode()


def test_boundary_settings_are_validated_on_the_host():
    """Boundary codes outside the supported set, or a physical code on a
    periodic direction, are rejected by the host set-up with a message."""
    from vcp3d_open_amd.deck import DeckCase
    from vcp3d_open_amd.solver import VCPError
    from tests.golden.make_golden import HERE as GOLDEN

    class Edited:
        def __init__(self, base, **kw):
            self.base, self.kw = base, kw
            self.rule, self.n_elems = base.rule, base.n_elems

        def description(self):
            return dict(self.base.description(), **self.kw)

        def element_grid(self):
            return self.base.element_grid()

    deck = DeckCase(os.path.join(GOLDEN, "ref_n3_8x4x4", "param_walls_a.cfg"))
    host_setup(deck)   # valid: subsonic inflow/outflow, isothermal walls
    with pytest.raises(VCPError, match="unsupported boundary condition"):
        host_setup(Edited(deck, bc=[12, 5, 7, 7, 0, 0]))
    with pytest.raises(VCPError, match="periodic direction"):
        host_setup(Edited(deck, bc=[4, 5, 7, 7, 6, 0]))
