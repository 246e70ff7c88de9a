"""Serving loop host logic (no GPU): the seeded request stream, the footprint
estimate's arithmetic and the latency summary."""
import math

import numpy as np

from vcp3d_open_amd.cases import OpenBoxCase, TGVCase
from vcp3d_open_amd.serve import Result, estimate_bytes, serve_workload, summary


def test_workload_is_seeded_and_poisson():
    a, b = serve_workload(seed=3, count=400), serve_workload(seed=3, count=400)
    assert [(r.kind, r.case.n_elem, r.case.p, r.arrival) for r in a] == \
           [(r.kind, r.case.n_elem, r.case.p, r.arrival) for r in b]
    gaps = np.diff([0.0] + [r.arrival for r in a])
    assert abs(gaps.mean() - 1 / 50) < 0.15 / 50              # rate 50 per second
    frac = np.mean([r.kind == "tgv" for r in a])
    assert 0.62 < frac < 0.78                                  # 70 / 30 mix
    assert all(r.rk_steps == 200 for r in a)
    assert {r.case.p for r in a if r.kind == "tgv"} == {2, 3, 4, 5, 6}
    boxes = [r.case for r in a if r.kind == "openbox"]
    assert all(0.3 <= c.mach <= 0.7 and c.n_elem == (48, 24, 24) and c.p == 4 for c in boxes)


def test_footprint_estimate_terms():
    # periodic TGV: per element 8 (10 M^3 + 66 M^2 + 16 + 15 K^3 + 60 K^2 + 2) + 28 bytes
    c = TGVCase(n=4, p=4)
    K, M = 5, 7
    per = 8 * (10 * M ** 3 + 66 * M ** 2 + 16 + 15 * K ** 3 + 60 * K ** 2 + 2) + 28
    est = estimate_bytes(c)
    assert 64 * per < est < 64 * per + 400_000
    # open box: boundary faces on iMin/iMax and the sponge elements add to it
    b0 = OpenBoxCase(n=(6, 2, 2), p=3, sponge_elems=0)
    b2 = OpenBoxCase(n=(6, 2, 2), p=3, sponge_elems=2)
    M3 = (b0.rule // 2 + 1) ** 3
    assert estimate_bytes(b2) - estimate_bytes(b0) == 4 * 24 + 8 * 8 * (8 + 13 * M3)


def test_summary_percentiles():
    res = [Result(i, "tgv", "done", arrival=0.1 * i, finished=0.1 * i + 1.0 + 0.01 * i, rk_steps=10, dofs=1000)
           for i in range(100)]
    res.append(Result(100, "tgv", "failed", arrival=0.0, finished=1.0))
    s = summary(res)
    assert s["done"] == 100 and s["failed"] == 1
    assert abs(s["p50_latency_s"] - (1.0 + 0.01 * 49.5)) < 1e-9
    assert math.isclose(s["gdof_updates"], 100 * 1000 * 3 * 10 / 1e9)
