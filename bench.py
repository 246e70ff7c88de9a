#!/usr/bin/env python3
"""Throughput of the gfx950 FP64 DG residual + SSP-RK3 step.

Workload: curvilinear Taylor-Green vortex, 32^3 hexahedra, P = 4
(K = 5 modes, M = 7 Gauss points per direction, rule 12), M = 0.1,
Re = 1600, nodes moved by x_i += 0.1 h sin x sin y sin z, FP64 throughout,
reference behaviour (metric indexing of the reference kept).
A "step" is one SSP-RK3 synchronization step (3 stages, each a full
residual evaluation with the per-element entropy mass-matrix solve).

value = 5 * nElem * (P+1)^3 * 3 stages * steps / seconds / 1e9
        (GDOF-updates/s per RK stage), inputs resident in HBM.

Also reported: the roofline position of the dominant kernel (its mean launch
time over the timed steps, from HIP events recorded around each launch on the
library's stream) and a CPU baseline:
one RK step of the same case by the reference itself, built from its sources
(oracle/_ref), on the host cores, with the C oracle's rate beside it (the
oracle alone where the reference build is absent or cannot run). Beside the
FP64 line: the mixed-precision form on the same case with its difference
from FP64 ("mixed"), and the batched (64 seeded Taylor-Green cases, FP64 and
mixed) and served (40-request Poisson stream, mixed, p50/p99 latency) forms
("forms"). --workload batch / serve / openbox measure those alone.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "GDOF-updates/s per RK stage, FP64 vs mixed at stated err; served p50/p99 latency"
PEAK_FP64_TFLOPS = 78.6   # MI355X dense FP64 (vector and matrix)
PEAK_HBM_GBS = 8000.0


def volume_kernel_flops(K: int, M: int, cg_iters: float) -> float:
    """Algorithmic FLOPs per element of the fused volume + mass-matrix kernel
    (2 per multiply-add; sum-factorised contraction counts, 350 flops per
    integration point for the pointwise flux work)."""
    fwd = 5 * (2 * K**3 * M + 3 * K**2 * M**2 + 4 * K * M**3)
    bwd = 5 * (3 * K * M**3 + 3 * K**2 * M**2 + 2 * K**3 * M)
    gather = 60 * K**3      # six faces x (value, derivative plane) x 5 variables
    cg = 5 * 2 * (K**3 * M + K**2 * M**2 + K * M**3) + 30 * M**3 + 25 * K**3 + 30 * K**3
    return 2.0 * (fwd + bwd + gather + cg_iters * cg) + 350.0 * M**3


def volume_kernel_bytes(K: int, M: int, update_share: float = 0.0) -> float:
    """Algorithmic HBM bytes per element of a residual launch: U, geometry
    (J + 9 metric terms per point), 16 element constants and the six face
    slots (value and normal-derivative planes, 10 K^2 each) read; R written.
    A launch with the fused RK-stage update also reads U0 (and writes U in
    place of R); update_share is the fraction of such launches averaged over
    (2/3 in an SSP-RK3 step)."""
    return 8.0 * (5 * K**3 + 60 * K**2 + 10 * M**3 + 16 + 5 * K**3 + update_share * 5 * K**3)


def measured_traffic(kernel: str):
    """HBM bytes per launch from the calibrated FETCH_SIZE/WRITE_SIZE passes
    recorded in profiles/ for this kernel, or None."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_tgv32_p4_fp64.json")
    try:
        with open(path) as f:
            data = json.load(f)
        return data["kernels"][kernel]["bytes"]
    except (OSError, KeyError, ValueError):
        return None


def face_kernel_flops(K: int, M: int) -> float:
    """Algorithmic FLOPs per element-face (both traces, flux, both lifts to
    the tangential coefficient planes); 700 flops per face point for the
    Ismail-Roe + viscous + penalty work."""
    trace = 2 * 5 * (2 * K**3 + 3 * K**2 * M + 4 * K * M**2)
    lift = 2 * 5 * (4 * K * M**2 + 3 * K**2 * M)
    return 2.0 * (trace + lift) + 700.0 * M**2


def face_kernel_bytes(K: int, M: int) -> float:
    return 8.0 * (2 * 5 * K**3 + 22 * M**2 + 2 * 10 * K**2 + 2)


def calib_copy(nbytes: int) -> int:
    """One device-to-device hipMemcpy of nbytes (the runtime's copy kernel):
    the reference point that scales the HBM byte counters."""
    import ctypes as C
    hip = C.CDLL(os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so"))
    src, dst = C.c_void_p(), C.c_void_p()
    for p in (src, dst):
        if hip.hipMalloc(C.byref(p), C.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(src, 1, C.c_size_t(nbytes))
    hip.hipDeviceSynchronize()
    if hip.hipMemcpy(dst, src, C.c_size_t(nbytes), 3) != 0:   # hipMemcpyDeviceToDevice
        raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(src)
    hip.hipFree(dst)
    return nbytes


def cpu_threads() -> int:
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


def port_baseline(case, label: str, steps: int = 1) -> dict:
    """C oracle (OpenMP over the host cores) on the same case: one RK step."""
    from oracle.pyoracle import Oracle, build
    build()
    threads = cpu_threads()
    o = Oracle(case)
    o.set_restart(case.initial_conservative())
    if hasattr(case, "damping_primitive") and getattr(case, "sponge_elems", 0) > 0:
        o.set_damping_state(case.damping_primitive())
    o.residual()
    t0 = time.perf_counter()
    for _ in range(steps):
        o.sync_step()
    dt = time.perf_counter() - t0
    gdof = 5 * case.n_elems * case.n_dofs * 3 * steps / dt / 1e9
    return {"value": gdof, "unit": "GDOF/s", "cores": threads, "kind": "port",
            "sample": f"{steps} SSP-RK3 step(s) of the same {label} case, C oracle, "
                      f"{threads} OpenMP threads, {dt:.2f} s"}


def reference_baseline(case, label: str):
    """The reference itself (VCP3D_CurviLinear built from its sources with its
    own Makefile flags into oracle/_ref/) on the same case, driven through its
    public API: the wall time of one ExecuteTimeSynchronizationStep (one RK
    step; its setup and input reading are not timed). Returns (result, None),
    or (None, reason) when the build is absent or does not run on this host."""
    import tempfile
    from oracle.pyoracle import REF_DRIVER, run_reference
    if not os.path.exists(REF_DRIVER):
        return None, "reference build not present"
    threads = cpu_threads()
    try:
        with tempfile.TemporaryDirectory(prefix="vcp_ref_") as d:
            prm = case.write_inputs(d)
            t0 = time.perf_counter()
            info = run_reference(prm, 1, (), threads=threads, timeout=600)
            wall = time.perf_counter() - t0
    except Exception as exc:  # a missing library or an unsupported instruction on this host
        return None, f"reference build did not run: {exc}"
    step = float(info["step_seconds"][0])
    gdof = 5 * case.n_elems * case.n_dofs * 3 / step / 1e9
    return {"value": gdof, "unit": "GDOF/s", "cores": threads, "kind": "reference",
            "sample": f"1 SSP-RK3 step (ExecuteTimeSynchronizationStep) of the same {label} case, "
                      f"reference VCP3D_CurviLinear built from its sources (-O2 -march=native, "
                      f"OpenMP), {threads} threads, {step:.2f} s per step ({wall:.1f} s with "
                      f"setup and I/O, untimed)"}, None


def cpu_baseline(case, workload: str) -> dict:
    """The reference build where it runs; the C oracle otherwise (and for the
    open box, where the reference's first step at P = 4 writes past a nodal
    buffer and aborts). The oracle's rate is reported beside it either way."""
    label = "48x24x24 P=4 open-box" if workload == "openbox" else "32^3 P=4"
    port = port_baseline(case, label)
    if workload == "openbox":
        return port
    ref, why = reference_baseline(case, label)
    if ref is None:
        port["note"] = why
        return port
    ref["port_value"] = port["value"]
    ref["port_sample"] = port["sample"]
    return ref


def bench_batch(args) -> int:
    """Config-3 workload: args.cases Taylor-Green cases (P uniform in 2..6, n^3
    elements with n uniform in [4, 24], Mach, Reynolds and phase drawn from
    one seed) stepped together by one batch: every kernel of an RK stage is
    one launch per (K, M) shape over all cases of that shape. value = the
    aggregate GDOF-updates/s per RK stage over all cases."""
    from vcp3d_open_amd import autotune
    from vcp3d_open_amd.cases import batch_workload
    from vcp3d_open_amd.solver import Batch, SolverClass

    t_setup = time.perf_counter()
    cases = batch_workload(seed=args.seed, count=args.cases)
    batch = Batch(use_graph=args.graph)
    members = []
    for c in cases:
        g = SolverClass(c, precision=args.precision)
        g.ReadRestartSolution(c.initial_conservative())
        batch.add(g, 0)
        members.append(g)
    t_setup = time.perf_counter() - t_setup
    print(f"[bench] {len(cases)} cases set up in {t_setup:.1f} s", file=sys.stderr, flush=True)
    if args.autotune:
        tuned = autotune.tune(batch, reps=args.profile_reps)
        autotune.save(tuned)
        tiles = {k: v["nt"] for k, v in tuned.items()}
    else:
        tiles = autotune.load_tiles()
        for (K, M), nt in tiles.items():
            if any((g["K"], g["M"]) == (K, M) for g in batch.groups()):
                batch.set_tile(K, M, nt)
    batch.step(args.warmup)
    batch.wait()
    print(f"[bench] warmup done ({args.warmup} steps)", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    batch.step(args.steps)
    batch.wait()
    secs = time.perf_counter() - t0
    print(f"[bench] timed {args.steps} batched steps in {secs:.3f} s", file=sys.stderr, flush=True)
    st = batch.status()
    bad = [i for i, m in enumerate(members) if st[m][3] != 0]
    if bad:
        raise RuntimeError(f"cases {bad} raised error flags")
    ndof = sum(5 * c.n_elems * c.n_dofs for c in cases)
    value = ndof * 3 * args.steps / secs / 1e9

    # dominant kernel: the volume launch of the shape group that takes longest
    groups = batch.groups()
    prof = batch.profile(args.profile_reps)
    per_group = []
    for g, (ms_face, ms_vol) in zip(groups, prof):
        K, M = g["K"], g["M"]
        mem = [m for m in members if (m.K, m.M) == (K, M)]
        cg = float(sum(m.cg_iterations().sum() for m in mem) / max(1, sum(m.n_elem for m in mem)))
        per_group.append({"P": K - 1, "K": K, "M": M, "cases": g["cases"], "elements": g["work"],
                          "volume_threads": g["nt"], "ms_face_kernel": ms_face, "ms_volume_kernel": ms_vol,
                          "cg_iters_mean": cg,
                          "volume_tflops": volume_kernel_flops(K, M, cg) * g["work"] / (ms_vol * 1e-3) / 1e12})
    top = max(per_group, key=lambda x: x["ms_volume_kernel"])
    K, M = top["K"], top["M"]
    flops = volume_kernel_flops(K, M, top["cg_iters_mean"]) * top["elements"]
    bytes_ = volume_kernel_bytes(K, M) * top["elements"]
    ms = top["ms_volume_kernel"]
    achieved = flops / (ms * 1e-3) / 1e12
    intensity = flops / bytes_
    attainable = min(PEAK_FP64_TFLOPS, PEAK_HBM_GBS * intensity / 1e3)
    roofline = {"bound": "valu-fp64" if attainable >= PEAK_FP64_TFLOPS else "hbm", "achieved": achieved,
                "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_TFLOPS,
                "traffic": None, "kernel": f"volume_kernel<{K}, {M}, {top['volume_threads']}, batch>",
                "ms_per_launch": ms, "algorithmic_flops": flops, "algorithmic_bytes": bytes_,
                "intensity_flop_per_byte": intensity, "attainable_tflops": attainable,
                "frac_of_attainable": achieved / attainable, "groups": per_group}
    out = {
        "metric": METRIC, "value": value, "unit": "GDOF/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": secs / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32-contractions/f64",
        "data": "synthetic",
        "config": {"workload": f"{len(cases)} batched Taylor-Green cases (seed {args.seed}): P uniform in "
                               "2..6, n^3 hex with n uniform in [4, 24], Mach U[0.05, 0.4], Re log-U[400, 3200], "
                               f"random phases; {args.precision}; one RK step of all cases per step",
                   "cases": len(cases), "n_elem": sum(c.n_elems for c in cases), "dofs": ndof,
                   "graph": bool(args.graph), "volume_threads": {f"{k},{m}": nt for (k, m), nt in tiles.items()},
                   "setup_s": t_setup},
        "roofline": roofline,
    }
    if not args.no_cpu_baseline:
        # the oracle on a bounded sample: one RK step of the smallest case of each degree
        sample = []
        for p in sorted({c.p for c in cases}):
            sample.append(min((c for c in cases if c.p == p), key=lambda c: c.n_elems))
        tot_dof, tot_s, threads = 0.0, 0.0, cpu_threads()
        from oracle.pyoracle import Oracle, build
        build()
        for c in sample:
            o = Oracle(c)
            o.set_restart(c.initial_conservative())
            o.residual()
            t1 = time.perf_counter()
            o.sync_step()
            tot_s += time.perf_counter() - t1
            tot_dof += 5 * c.n_elems * c.n_dofs * 3
        out["cpu_baseline"] = {"value": tot_dof / tot_s / 1e9, "unit": "GDOF/s", "cores": threads, "kind": "port",
                               "sample": "1 SSP-RK3 step of the smallest case of each degree of the batch ("
                                         + ", ".join(f"P={c.p} n={c.n}" for c in sample)
                                         + f"), C oracle, {threads} OpenMP threads, {tot_s:.2f} s"}
    print(json.dumps(out))
    return 0


def forms_summary(args) -> dict:
    """The batched (config 3) and served (config 4) forms beside the default
    line: the 64-case batch in FP64 and in mixed precision (the same cases,
    10 timed RK steps after 2 warm-up steps each), and the 40-request served
    stream in mixed precision."""
    from vcp3d_open_amd import autotune, serve
    from vcp3d_open_amd.cases import batch_workload
    from vcp3d_open_amd.solver import Batch, SolverClass

    cases = batch_workload(seed=2024, count=64)
    solvers = []
    for c in cases:
        g = SolverClass(c)
        g.ReadRestartSolution(c.initial_conservative())
        solvers.append(g)
    ndof = sum(5 * c.n_elems * c.n_dofs for c in cases)
    out = {"batch": {"cases": len(cases), "dofs": ndof, "rk_steps_timed": 10}}
    for precision in ("fp64", "mixed"):
        for g in solvers:
            g.set_precision(precision)
        b = Batch(tiles=autotune.load_tiles())
        for g in solvers:
            b.add(g, 0)
        b.step(2)
        b.wait()
        t0 = time.perf_counter()
        b.step(10)
        b.wait()
        secs = time.perf_counter() - t0
        flags = [int(w[3]) for w in b.status().values()]
        b.close()
        out["batch"][precision] = {"value": ndof * 3 * 10 / secs / 1e9, "unit": "GDOF/s",
                                   "ms_per_rk_step": secs / 10 * 1e3, "cases_with_error_flags": sum(f != 0 for f in flags)}
        print(f"[bench] batch {precision}: {out['batch'][precision]['value']:.2f} GDOF/s", file=sys.stderr, flush=True)
    out["batch"]["mixed_speedup"] = out["batch"]["mixed"]["value"] / out["batch"]["fp64"]["value"]
    for g in solvers:
        g.close()
    reqs = serve.serve_workload(seed=2024, count=40)
    srv = serve.Server(precision="mixed", tiles=autotune.load_tiles())
    res = srv.run(reqs)
    summ = serve.summary(res)
    out["serve"] = {"precision": "mixed", "requests": len(reqs), "rate_per_s": 50, "rk_steps": 200,
                    "value": summ["gdof_per_s"], "unit": "GDOF/s",
                    "p50_latency_s": summ["p50_latency_s"], "p99_latency_s": summ["p99_latency_s"],
                    "done": summ["done"], "failed": summ["failed"], "makespan_s": summ["makespan_s"],
                    "peak_concurrent": srv.peak_active, "peak_device_bytes": srv.peak_bytes}
    print(f"[bench] serve mixed: {summ['gdof_per_s']:.2f} GDOF/s, p50 {summ['p50_latency_s']:.2f} s, "
          f"p99 {summ['p99_latency_s']:.2f} s", file=sys.stderr, flush=True)
    return out


def bench_serve(args) -> int:
    """Config-4 workload: args.cases requests arriving as a Poisson stream
    (50 per second, seeded), 70 % Taylor-Green cases of the batched workload's
    distribution and 30 % 48x24x24 P=4 open boxes (PC-NSCBC, sponge + CML,
    Mach U[0.3, 0.7]), 200 RK steps each, served by continuous batching with
    admission by device footprint (80 % of 288 GB). value = aggregate
    GDOF-updates/s per RK stage of the completed requests over the makespan
    (first arrival to last completion); p50/p99 request latency (arrival to
    completion, set-up included) beside it."""
    import numpy as np
    from vcp3d_open_amd import autotune, serve
    from vcp3d_open_amd.solver import Batch, SolverClass

    reqs = serve.serve_workload(seed=args.seed, count=args.cases, rk_steps=args.rk_steps)
    # per-request error of the mixed form against FP64: the three smallest requests
    sample = sorted(reqs, key=lambda r: r.dofs)[:3] if args.precision == "mixed" else []
    srv = serve.Server(precision=args.precision, steps_per_tick=args.tick, workers=4,
                       keep_solutions={r.rid for r in sample}, tiles=autotune.load_tiles())
    t0 = time.perf_counter()
    res = srv.run(reqs)
    wall = time.perf_counter() - t0
    summ = serve.summary(res)
    print(f"[bench] served {summ['done']}/{len(reqs)} requests in {wall:.1f} s", file=sys.stderr, flush=True)
    errors = []
    for r in sample:
        s = SolverClass(r.case)
        s.ReadRestartSolution(r.case.initial_conservative())
        if getattr(r.case, "sponge_elems", 0):
            s.SetDampingState(r.case.damping_primitive())
        b = Batch()
        b.add(s, r.rk_steps)
        b.step(r.rk_steps)
        b.wait()
        ref = s.solution()
        got = next(x for x in res if x.rid == r.rid).solution
        errors.append({"rid": r.rid, "kind": r.kind, "P": r.case.p, "n_elem": list(r.case.n_elem),
                       "modal_rel_l2_vs_fp64": float(np.linalg.norm(got - ref) / np.linalg.norm(ref))})
        b.close()
        s.close()
    out = {
        "metric": METRIC, "value": summ["gdof_per_s"], "unit": "GDOF/s", "n_gpus": 1,
        "steps": args.rk_steps, "warmup": 0, "ms_per_step": None,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32-contractions/f64",
        "data": "synthetic",
        "config": {"workload": f"served Poisson stream (seed {args.seed}, 50 requests/s) of {args.cases} requests: "
                               "70 % Taylor-Green (P 2..6, n in [4, 24]), 30 % open box 48x24x24 P=4 "
                               f"(M in [0.3, 0.7]), {args.rk_steps} RK steps each; continuous batching, "
                               "admission by device footprint <= 80 % of 288 GB",
                   "requests": args.cases, "steps_per_tick": args.tick},
        "serving": {**summ, "peak_device_bytes": srv.peak_bytes, "peak_concurrent": srv.peak_active,
                    "ticks": srv.ticks, "wall_s": wall,
                    "latency_s": {str(r.rid): round(r.latency, 4) for r in res if r.status == "done"}},
        "error_vs_fp64_sample": errors,
    }
    print(json.dumps(out))
    return 0


def bench_interpolation(args) -> int:
    """The mass-matrix solve's forward interpolation (5 variables, K^3 -> M^3,
    K = 5, M = 7) over 32768 synthetic elements, 20 times per launch, as FP64
    VALU pencils (the volume kernel's form) and as 16x16x4 FP64 MFMA tiles
    (libvcp3d_ab.so); both launch times, the rate in useful FLOP/s and the
    relative difference of the two outputs."""
    import ctypes as C
    import os
    lib = C.CDLL(os.path.join(os.path.dirname(os.path.abspath(__file__)), "vcp3d_open_amd", "libvcp3d_ab.so"))
    lib.vcp_ab_interpolation.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double),
                                         C.POINTER(C.c_double)]
    n_elem, inner, reps = 32768, 20, max(args.steps, 3)
    K, M = 5, 7
    flop = 2.0 * 5 * (K * K * K * M + K * K * M * M + K * M * M * M) * n_elem * inner   # useful
    res = {}
    for impl, name in ((0, "valu"), (1, "mfma")):
        ms, tot = C.c_double(), C.c_double()
        rc = lib.vcp_ab_interpolation(impl, n_elem, inner, reps, C.byref(ms), C.byref(tot))
        if rc != 0:
            print(json.dumps({"error": f"vcp_ab_interpolation({name}) failed: {rc}"}))
            return 1
        res[name] = (ms.value, tot.value)
    rate = {k: flop / (v[0] * 1e-3) / 1e12 for k, v in res.items()}
    out = {"metric": "mass_matrix_interpolation_tflops", "value": rate["valu"], "unit": "TFLOP/s",
           "higher_is_better": True, "n_gpus": 1, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"forward K^3->M^3 interpolation of 5 variables, K=5, M=7, {n_elem} elements "
                                  f"x {inner} per launch, {reps} launches"},
           "valu_ms": res["valu"][0], "mfma_ms": res["mfma"][0],
           "valu_tflops": rate["valu"], "mfma_tflops": rate["mfma"],
           "peak_fp64_tflops": PEAK_FP64_TFLOPS,
           "rel_diff": abs(res["mfma"][1] - res["valu"][1]) / abs(res["valu"][1])}
    print(json.dumps(out))
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=32)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--profile-reps", type=int, default=5)
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="also run one device-to-device copy of this size (HBM counter calibration)")
    ap.add_argument("--rk-steps", type=int, default=200, help="serve: RK steps per request")
    ap.add_argument("--tick", type=int, default=8, help="serve: batched RK steps between status reads")
    ap.add_argument("--workload", choices=["tgv", "openbox", "batch", "serve", "interpolation"], default="tgv",
                    help="tgv: curvilinear Taylor-Green 32^3 P=4 (default); openbox: 48x24x24 P=4 "
                         "PC-NSCBC in/outflow, 8-element sponge + CML, M=0.5; batch: 64 seeded "
                         "Taylor-Green cases of uneven P and size stepped together")
    ap.add_argument("--seed", type=int, default=2024, help="batch: seed of the case draw")
    ap.add_argument("--cases", type=int, default=64, help="batch: number of cases; serve: number of requests")
    ap.add_argument("--autotune", action="store_true",
                    help="batch: time every volume workgroup size per shape and record the table")
    ap.add_argument("--graph", action="store_true", help="batch: replay each RK step as a HIP graph")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64",
                    help="arithmetic of the measured line: fp64 (default) or mixed (FP32 contractions and "
                         "mass-matrix solve, FP64 state, pointwise physics, metric, boundaries and RK update)")
    ap.add_argument("--no-mixed", action="store_true",
                    help="tgv/openbox: skip the mixed-precision measurement beside the FP64 line")
    ap.add_argument("--no-forms", action="store_true",
                    help="tgv: skip the batched and served forms beside the default line")
    ap.add_argument("--accuracy-steps", type=int, default=0,
                    help="tgv/openbox: also advance FP64 and mixed this many RK steps from the same state "
                         "and report their field, kinetic-energy and enstrophy differences")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        return 2
    if args.workload == "batch":
        return bench_batch(args)
    if args.workload == "serve":
        return bench_serve(args)
    if args.workload == "interpolation":
        return bench_interpolation(args)

    from vcp3d_open_amd.cases import OpenBoxCase, TGVCase
    from vcp3d_open_amd.solver import SolverClass

    if args.workload == "openbox":
        case = OpenBoxCase(n=(48, 24, 24), p=4, rule=12, sponge_elems=8)
        workload = ("open box 48x24x24 hex, P=4, rule 12, PC-NSCBC static inflow/outflow, "
                    "8-element sponge + CML at the outlet, M=0.5, vortex + pressure pulse, FP64")
    else:
        case = TGVCase(n=args.n, p=4, rule=12, curvilinear=True)
        workload = f"curvilinear Taylor-Green vortex {args.n}^3 hex, P=4, rule 12, FP64"
    def make(precision):
        s_ = SolverClass(case, precision=precision)
        s_.ReadRestartSolution(case.initial_conservative())
        if args.workload == "openbox":
            s_.SetDampingState(case.damping_primitive())
        s_.Residual(True)
        return s_

    def timed(s_, label):
        for _ in range(args.warmup):
            s_.ExecuteTimeSynchronizationStep()
        print(f"[bench] {label}: warmup done ({args.warmup} steps)", file=sys.stderr, flush=True)
        # per-launch durations of the timed steps' face and volume kernels
        # (HIP events on the library stream, around each launch)
        s_.set_kernel_timing(True)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            md_ = s_.ExecuteTimeSynchronizationStep()
        secs_ = time.perf_counter() - t0
        kt_ = s_.kernel_timing()
        s_.set_kernel_timing(False)
        print(f"[bench] {label}: timed {args.steps} steps in {secs_:.3f} s", file=sys.stderr, flush=True)
        return secs_, md_, kt_

    solver = make(args.precision)
    secs, md, ktimes = timed(solver, args.precision)
    ndof = 5 * case.n_elems * case.n_dofs
    value = ndof * 3 * args.steps / secs / 1e9

    # dominant kernel: mean per-launch times of the timed steps' launches
    # (HIP events on the library stream), with residual-only launches of the
    # same kernels beside them
    cg = solver.cg_iterations()
    cg_mean = float(cg.mean())
    ms_face, ms_vol, n_timed = ktimes
    rs_face, rs_vol = solver.profile_residual(args.profile_reps)
    K, M = case.p + 1, 7
    if ms_vol >= ms_face:
        name, ms = "volume_kernel<5, 7, 256>", ms_vol
        flops = volume_kernel_flops(K, M, cg_mean) * case.n_elems
        bytes_ = volume_kernel_bytes(K, M, 2.0 / 3.0) * case.n_elems
    else:
        name, ms = "face_kernel<5, 7, BUG> (3 directions)", ms_face
        flops = face_kernel_flops(K, M) * 3 * case.n_elems
        bytes_ = face_kernel_bytes(K, M) * 3 * case.n_elems
    achieved = flops / (ms * 1e-3) / 1e12
    # the binding roof: min(FP64 peak, HBM bandwidth x algorithmic intensity)
    intensity = flops / bytes_
    attainable = min(PEAK_FP64_TFLOPS, PEAK_HBM_GBS * intensity / 1e3)
    hbm_gbs = bytes_ / (ms * 1e-3) / 1e9
    # the compute roof is the FP64 VALU (vector FMA) peak, which MI355X's FP64
    # MFMA peak equals; no MFMA instruction runs in these kernels
    roofline = {"bound": "valu-fp64" if attainable >= PEAK_FP64_TFLOPS else "hbm", "achieved": achieved,
                "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / PEAK_FP64_TFLOPS, "traffic": measured_traffic(name) if args.workload == "tgv" else None, "kernel": name,
                "ms_per_launch": ms, "algorithmic_flops": flops, "algorithmic_bytes": bytes_,
                "intensity_flop_per_byte": intensity, "attainable_tflops": attainable,
                "frac_of_attainable": achieved / attainable,
                "hbm_gbs_algorithmic": hbm_gbs, "hbm_frac": hbm_gbs / PEAK_HBM_GBS,
                "ms_face_kernel": ms_face, "ms_volume_kernel": ms_vol, "timed_stages": n_timed,
                "ms_face_kernel_residual_only": rs_face, "ms_volume_kernel_residual_only": rs_vol,
                "cg_iters_mean": cg_mean}

    out = {
        "metric": METRIC, "value": value, "unit": "GDOF/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": secs / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32-contractions/f64",
        "data": "synthetic",
        "config": {"workload": workload,
                   "n_elem": case.n_elems, "poly_degree": 4, "dofs": ndof, "mach": case.mach,
                   "reynolds": case.reynolds, "dt_synchr": case.dt_synchr},
        "roofline": roofline,
        "mach2_max": float(md[6]),
    }
    if args.precision == "fp64" and not args.no_mixed:
        # the mixed-precision form on the same case, same steps, beside the FP64 line,
        # with its difference from the FP64 state after the same steps
        import numpy as np
        from vcp3d_open_amd.diagnostics import flow_integrals
        from vcp3d_open_amd.solver import host_setup
        mixed = make("mixed")
        msecs, _, mkt = timed(mixed, "mixed")
        mvalue = ndof * 3 * args.steps / msecs / 1e9
        mf_face, mf_vol, _ = mkt
        out["mixed"] = {"value": mvalue, "unit": "GDOF/s", "ms_per_step": msecs / args.steps * 1e3,
                        "speedup_vs_fp64": mvalue / value, "dtype": "f32-contractions/f64",
                        "ms_face_kernel": mf_face, "ms_volume_kernel": mf_vol,
                        "cg_iters_mean": float(mixed.cg_iterations().mean())}
        n_acc = args.accuracy_steps
        ref_s, mix_s = solver, mixed
        if n_acc:
            ref_s, mix_s = make("fp64"), make("mixed")
            for s_ in (ref_s, mix_s):
                done = 0
                while done < n_acc:
                    s_.ExecuteTimeSynchronizationStep()
                    done += s_.last_rk_steps
        na, nb = mix_s.WriteRestartSolution(), ref_s.WriteRestartSolution()
        vol = host_setup(case)["vol"]
        fm, fo = flow_integrals(case, mix_s.solution(), vol), flow_integrals(case, ref_s.solution(), vol)
        out["mixed"]["error_vs_fp64"] = {
            "rk_steps": n_acc or args.warmup + args.steps,
            "field_rel_l2": float(np.linalg.norm(na - nb) / np.linalg.norm(nb)),
            "field_rel_l2_per_variable": [float(np.linalg.norm(na[..., v] - nb[..., v]) / np.linalg.norm(nb[..., v]))
                                          for v in range(5)],
            "kinetic_energy_rel": abs(fm["kinetic_energy"] / fo["kinetic_energy"] - 1),
            "enstrophy_rel": abs(fm["enstrophy"] / fo["enstrophy"] - 1)}
    if args.workload == "tgv" and args.precision == "fp64" and not args.no_forms:
        out["forms"] = forms_summary(args)
    if args.calib_copy_mib:
        out["calib_copy_bytes"] = calib_copy(args.calib_copy_mib << 20)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(case, args.workload)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
