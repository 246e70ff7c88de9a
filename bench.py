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
time measured with HIP events on the library's stream) and a CPU baseline
(the C oracle, OpenMP over the host cores, one RK step of the same case).
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


def volume_kernel_bytes(K: int, M: int) -> float:
    """Algorithmic HBM bytes per element of a residual launch: U, geometry
    (J + 9 metric terms per point), 16 element constants and the six face
    slots (value and normal-derivative planes, 10 K^2 each) read; R written."""
    return 8.0 * (5 * K**3 + 60 * K**2 + 10 * M**3 + 16 + 5 * K**3)


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


def cpu_baseline(case, steps: int = 1) -> dict:
    """C oracle (OpenMP over the host cores) on the same case: one RK step."""
    from oracle.pyoracle import Oracle, build
    build()
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
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
            "sample": f"{steps} SSP-RK3 step(s) of the same 32^3 P=4 case, C oracle, "
                      f"{threads} OpenMP threads, {dt:.2f} s"}


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
    ap.add_argument("--workload", choices=["tgv", "openbox"], default="tgv",
                    help="tgv: curvilinear Taylor-Green 32^3 P=4 (default); openbox: 48x24x24 P=4 "
                         "PC-NSCBC in/outflow, 8-element sponge + CML, M=0.5")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        return 2

    from vcp3d_open_amd.cases import OpenBoxCase, TGVCase
    from vcp3d_open_amd.solver import SolverClass

    if args.workload == "openbox":
        case = OpenBoxCase(n=(48, 24, 24), p=4, rule=12, sponge_elems=8)
        workload = ("open box 48x24x24 hex, P=4, rule 12, PC-NSCBC static inflow/outflow, "
                    "8-element sponge + CML at the outlet, M=0.5, vortex + pressure pulse, FP64")
    else:
        case = TGVCase(n=args.n, p=4, rule=12, curvilinear=True)
        workload = f"curvilinear Taylor-Green vortex {args.n}^3 hex, P=4, rule 12, FP64"
    solver = SolverClass(case)
    solver.ReadRestartSolution(case.initial_conservative())
    if args.workload == "openbox":
        solver.SetDampingState(case.damping_primitive())
    solver.Residual(True)
    for _ in range(args.warmup):
        solver.ExecuteTimeSynchronizationStep()
    print(f"[bench] warmup done ({args.warmup} steps)", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        md = solver.ExecuteTimeSynchronizationStep()
    t1 = time.perf_counter()
    secs = t1 - t0
    ndof = 5 * case.n_elems * case.n_dofs
    value = ndof * 3 * args.steps / secs / 1e9
    print(f"[bench] timed {args.steps} steps in {secs:.3f} s", file=sys.stderr, flush=True)

    # dominant kernel: per-launch times from HIP events on the library stream
    cg = solver.cg_iterations()
    cg_mean = float(cg.mean())
    ms_face, ms_vol = solver.profile_residual(args.profile_reps)
    K, M = case.p + 1, 7
    if ms_vol >= ms_face:
        name, ms = "volume_kernel<5, 7, 256>", ms_vol
        flops = volume_kernel_flops(K, M, cg_mean) * case.n_elems
        bytes_ = volume_kernel_bytes(K, M) * case.n_elems
    else:
        name, ms = "face_kernel<5, 7, BUG> (3 directions)", ms_face
        flops = face_kernel_flops(K, M) * 3 * case.n_elems
        bytes_ = face_kernel_bytes(K, M) * 3 * case.n_elems
    achieved = flops / (ms * 1e-3) / 1e12
    roofline = {"bound": "mfma", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / PEAK_FP64_TFLOPS, "traffic": measured_traffic(name) if args.workload == "tgv" else None, "kernel": name,
                "ms_per_launch": ms, "algorithmic_flops": flops, "algorithmic_bytes": bytes_,
                "hbm_gbs_algorithmic": bytes_ / (ms * 1e-3) / 1e9,
                "ms_face_kernel": ms_face, "ms_volume_kernel": ms_vol, "cg_iters_mean": cg_mean}

    out = {
        "metric": METRIC, "value": value, "unit": "GDOF/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": secs / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": workload,
                   "n_elem": case.n_elems, "poly_degree": 4, "dofs": ndof, "mach": case.mach,
                   "reynolds": case.reynolds, "dt_synchr": case.dt_synchr},
        "roofline": roofline,
        "mach2_max": float(md[6]),
    }
    if args.calib_copy_mib:
        out["calib_copy_bytes"] = calib_copy(args.calib_copy_mib << 20)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(case)
        if args.workload == "openbox":
            out["cpu_baseline"]["sample"] = out["cpu_baseline"]["sample"].replace(
                "32^3 P=4", "48x24x24 P=4 open-box")
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
