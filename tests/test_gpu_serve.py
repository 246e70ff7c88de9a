"""Serving loop on the GPU (continuous batching over vcp_batch): requests
arriving while others run join the batch between ticks, and each request's
result is bit-identical to the same request run alone (a batch of one), in
FP64 and in the mixed-precision form; the mixed results stay within 1e-7 of
FP64 after these few steps. A request whose state is invalid is evicted and
reported failed without changing the others. The footprint estimate used for
admission matches the contexts' own count."""
import numpy as np
import pytest

from vcp3d_open_amd.cases import OpenBoxCase, TGVCase
from vcp3d_open_amd.serve import Request, Server, estimate_bytes

pytestmark = pytest.mark.gpu


class BadTGV(TGVCase):
    """A Taylor-Green case whose initial state has a negative density in one
    element (no valid entropy variables there)."""

    def initial_conservative(self, coords=None):
        u = super().initial_conservative(coords)
        u[3, :, 0] = -1.0
        return u


def requests():
    kws = [dict(n=3, p=2, mach=0.3), dict(n=4, p=3, curvilinear=True), dict(n=2, p=4), dict(n=4, p=5, curvilinear=True),
           dict(n=2, p=6, mach=0.2)]
    # the first request runs long enough (400 RK steps) for the others to be
    # built and join while it is on the GPU
    out = [Request(i, TGVCase(**kw), 400 if i == 0 else 6 + 2 * i, 0.01 * i) for i, kw in enumerate(kws)]
    out.append(Request(5, OpenBoxCase(n=(6, 2, 2), p=3, sponge_elems=2), 8, 0.02, "openbox"))
    return out


def alone(req, precision):
    from vcp3d_open_amd.solver import Batch, SolverClass
    s = SolverClass(req.case, precision=precision)
    s.ReadRestartSolution(req.case.initial_conservative())
    if getattr(req.case, "sponge_elems", 0):
        s.SetDampingState(req.case.damping_primitive())
    b = Batch()
    b.add(s, req.rk_steps)
    b.step(req.rk_steps)
    b.wait()
    out = s.solution()
    b.close()
    s.close()
    return out


def test_served_results_equal_alone():
    sols = {}
    for precision in ("fp64", "mixed"):
        print(f"[case] serving 6 requests, {precision}", flush=True)
        srv = Server(precision=precision, steps_per_tick=2, workers=2, keep_solutions=True)
        res = srv.run(requests())
        assert all(r.status == "done" for r in res), [(r.rid, r.status, r.message) for r in res]
        assert srv.ticks > 2 and srv.peak_active >= 2          # requests overlapped on the GPU
        for r, req in zip(res, requests()):
            assert r.rk_steps == req.rk_steps
            assert np.array_equal(r.solution, alone(req, precision)), (precision, r.rid)
        sols[precision] = {r.rid: r.solution for r in res}
    for rid in sols["fp64"]:
        a, b = sols["mixed"][rid], sols["fp64"][rid]
        assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-7, rid


def test_failing_request_is_evicted():
    reqs = requests()[:3]
    reqs.append(Request(9, BadTGV(n=3, p=3), 6, 0.0))
    print("[case] serving with one invalid request", flush=True)
    srv = Server(steps_per_tick=2, workers=2, keep_solutions=True)
    res = {r.rid: r for r in srv.run(reqs)}
    assert res[9].status == "failed" and res[9].flags != 0
    for req in reqs[:3]:
        assert res[req.rid].status == "done"
        assert np.array_equal(res[req.rid].solution, alone(req, "fp64"))


def test_footprint_estimate_matches_the_context():
    from vcp3d_open_amd.solver import SolverClass
    for case in (TGVCase(n=4, p=4, curvilinear=True), TGVCase(n=3, p=6), OpenBoxCase(n=(6, 2, 2), p=3, sponge_elems=2)):
        s = SolverClass(case)
        actual, est = s.device_bytes(), estimate_bytes(case)
        s.close()
        print(f"[case] footprint {type(case).__name__} {case.n_elem} P={case.p}: {actual} vs {est}", flush=True)
        assert actual <= est <= actual + 400_000
