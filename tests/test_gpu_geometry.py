"""The metric terms vcp_create computes on the device (geometry.hip) against
the host set-up (vcp_host_setup, the same arithmetic on the CPU), through the
C ABI: volume metric, face slots and length scales bit for bit, averaged
normals against the host set-up's sequential sums, the maps by construction.
The host set-up itself is pinned against the oracle (test_host_setup.py)."""
import os
import time

import numpy as np
import pytest

from vcp3d_open_amd.cases import OpenBoxCase, TGVCase

pytestmark = pytest.mark.gpu

CASES = [
    ("tgv_curv_n4_p4", lambda: TGVCase(n=4, p=4, curvilinear=True)),
    ("tgv_cart_n3_p2", lambda: TGVCase(n=3, p=2)),
    ("tgv_curv_n4_p6", lambda: TGVCase(n=4, p=6, rule=16, curvilinear=True)),
    ("tgv_curv_n4_p3_wale", lambda: TGVCase(n=4, p=3, curvilinear=True, sgs="wale")),
    ("openbox_6x3x3_p3", lambda: OpenBoxCase(n=(6, 3, 3), p=3, sponge_elems=2)),
]


@pytest.fixture(scope="module")
def lib():
    from vcp3d_open_amd.solver import load_library
    return load_library()


def host_expectation(case, setup):
    """Device layouts assembled from the host set-up: vgeo = vol[:, :10];
    fgeo per (element, direction) = the neighbour below's max-face normal,
    area and inverse metric (the element's own min face where there is no
    neighbour) and the element's min-face inverse metric; the length scales and
    averaged normals summed point by point in the host set-up's order."""
    vol, faces, lens, nbr = setup["vol"], setup["faces"], setup["lens"], setup["nbr"]
    E = vol.shape[0]
    n2 = faces.shape[-1]
    fgeo = np.zeros((E, 3, 22, n2))
    for e in range(E):
        for d in range(3):
            L = nbr[e, 2 * d]
            mR = faces[2 * d, e]
            mL = faces[2 * d + 1, L] if L >= 0 else mR
            fgeo[e, d, 0:4] = mL[0:4]
            fgeo[e, d, 4:13] = mL[4:13]
            fgeo[e, d, 13:22] = mR[4:13]
    M = int(round(n2 ** 0.5))
    w, _ = None, None
    r, wq = np.polynomial.legendre.leggauss(M)
    w2 = [float(wq[j] * wq[k]) for k in range(M) for j in range(M)]
    ave = np.zeros((E, 9))
    for e in range(E):
        for d in range(3):
            s0, s1 = faces[2 * d, e], faces[2 * d + 1, e]
            an = [0.0, 0.0, 0.0]
            area = 0.0
            for l in range(n2):
                ww = w2[l]
                for c in range(3):
                    an[c] += ww * (float(s0[c, l]) * float(s0[3, l]) + float(s1[c, l]) * float(s1[3, l]))
                area += ww * (float(s0[3, l]) + float(s1[3, l]))
            for c in range(3):
                ave[e, 3 * d + c] = an[c] / area
    return vol[:, :10], fgeo, lens, ave


@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
def test_device_metrics_equal_host_setup(lib, name, make):
    from vcp3d_open_amd.solver import SolverClass, host_setup
    case = make()
    print(f"[case] device geometry {name}", flush=True)
    s = SolverClass(case)
    dev = s.device_geometry()
    vg, fg, lens, ave = host_expectation(case, host_setup(case))
    assert np.array_equal(dev["vgeo"], vg), np.abs(dev["vgeo"] - vg).max()
    assert np.array_equal(dev["fgeo"], fg), np.abs(dev["fgeo"] - fg).max()
    assert np.array_equal(dev["econst"][:, 0:4], lens), np.abs(dev["econst"][:, 0:4] - lens).max()
    # the host set-up's sums in numpy-free Python order; the host's own
    # weights come from its Gauss rule, equal to numpy's to rounding
    np.testing.assert_allclose(dev["econst"][:, 4:13], ave, rtol=0, atol=1e-14)
    if getattr(case, "sgs", "none") != "none":
        les = np.cbrt(lens[:, 3]) / case.p
        np.testing.assert_allclose(dev["econst"][:, 13], les, rtol=1e-15)
    s.close()


def test_creation_time_32cubed(lib):
    """Context creation of the 32^3 P=4 curvilinear case with the metric terms
    on the device (printed; the host set-up took 1.17 s on 16 threads)."""
    from vcp3d_open_amd.solver import SolverClass
    case = TGVCase(n=32, p=4, rule=12, curvilinear=True)
    grid = case.element_grid()
    t0 = time.perf_counter()
    s = SolverClass(case, elem_grid=grid)
    dt = time.perf_counter() - t0
    print(f"[case] 32^3 P=4 context creation {dt:.3f} s", flush=True)
    assert s.device_geometry()["econst"][:, 3].min() > 0.0
    s.close()
