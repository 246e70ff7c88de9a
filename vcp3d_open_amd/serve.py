"""Serving loop: continuous batching of solve requests on one GPU.

The reference runs one case per process (VCP3D_CurviLinear.cpp main, one
SolverClass) and stops the process on any error (Terminate,
StandAloneFunctions.cpp:628-697; the CG abort, CG_MassMatrix.cpp:398-400).
Here a request is one case advanced a given number of RK steps. The loop:
- admits arrived requests in arrival order while the summed device
  footprint of the admitted ones stays within the capacity (estimate_bytes
  before creation, the context's own count, vcp_device_bytes, after);
- builds the admitted cases (grid, initial state, set-up and upload) on a
  pool of host threads, so the GPU keeps stepping the running ones;
- adds each built case to one vcp_batch as soon as it is ready: it joins the
  running batch between two ticks of steps_per_tick batched RK steps;
- after every tick reads the batch status once; a case that has taken its
  steps leaves the batch and frees its memory, and a case whose error flags
  are set (non-convergent solve, a^2 negative or not a number, non-finite
  residual) is deactivated on the device by the batch itself, leaves and is
  reported as failed while the others continue.
A request's result is the same alone (a batch of one) and in the mix: the
batch kernels treat every element of every case independently.
"""
from __future__ import annotations

import collections
import concurrent.futures as cf
import math
import time
from dataclasses import dataclass, field

import numpy as np

from .cases import BATCH_RULE, OpenBoxCase, TGVCase
from .solver import ACTIVE, FLAGS, STEPS_DONE, Batch, SolverClass

HBM_BYTES = 288 * 10 ** 9          # one MI355X


def estimate_bytes(case) -> int:
    """Device bytes a context of this case allocates (vcp_create): geometry,
    state, face slots and per-element scalars, boundary faces and sponge
    arrays, plus an upper bound for the small table block."""
    K, M = case.p + 1, case.rule // 2 + 1
    ni, nj, nk = case.n_elem
    E = ni * nj * nk
    n3, n2, nd = M ** 3, M * M, K ** 3
    dbl = 10 * n3 + 3 * 22 * n2 + 16 + 15 * nd + 60 * K * K + 2
    per_elem = 8 * dbl + 4 * (6 + 1)
    desc = case.description()
    periodic = desc.get("periodic", (1, 1, 1))
    n_bface = sum(2 * (nj * nk, ni * nk, ni * nj)[d] for d in range(3) if not periodic[d])
    per_bface = 4 * 2 + 8 * (13 * n2 + 1 + 6)
    n_sp = 0
    if desc.get("sponge_layer"):
        se = desc.get("sponge_elems", (0,) * 6)
        i, j, k = np.meshgrid(np.arange(ni), np.arange(nj), np.arange(nk), indexing="ij")
        inside = ((i < se[0]) | (i >= ni - se[1]) | (j < se[2]) | (j >= nj - se[3])
                  | (k < se[4]) | (k >= nk - se[5]))
        n_sp = int(inside.sum())
    sponge = (4 * E + 8 * n_sp * (8 + 13 * n3)) if desc.get("sponge_layer") else 0
    avg = 8 * E * 11 * nd if desc.get("compute_average") else 0
    tables = 8 * (6 * M * K + K * K + 8 * K + n3 + 2 * n2 + n2 * n2 + 64) + 8 * 16 + 4 * 4
    return E * per_elem + n_bface * per_bface + sponge + avg + tables


@dataclass
class Request:
    rid: int
    case: object
    rk_steps: int
    arrival: float                  # seconds after the start of the run
    kind: str = "tgv"

    @property
    def dofs(self) -> int:
        return 5 * self.case.n_elems * self.case.n_dofs


@dataclass
class Result:
    rid: int
    kind: str
    status: str                     # "done", "failed" or "rejected"
    arrival: float
    admitted: float = math.nan      # admission (build starts)
    started: float = math.nan       # joined the batch
    finished: float = math.nan
    rk_steps: int = 0
    dofs: int = 0
    flags: int = 0
    device_bytes: int = 0
    message: str = ""
    solution: np.ndarray | None = field(default=None, repr=False)

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


class Server:
    """Continuous-batching request loop over one vcp_batch (one GPU)."""

    def __init__(self, capacity_bytes: int = int(0.8 * HBM_BYTES), precision: str = "fp64",
                 steps_per_tick: int = 8, workers: int = 4, keep_solutions=False,
                 tiles: dict | None = None):
        """keep_solutions: True, or a set of request ids, whose final modal
        states are copied into their results."""
        self.capacity = int(capacity_bytes)
        self.precision = precision
        self.steps_per_tick = int(steps_per_tick)
        self.workers = int(workers)
        self.keep_solutions = keep_solutions
        self.tiles = dict(tiles or {})
        self.peak_bytes = 0
        self.peak_active = 0
        self.ticks = 0

    def _build(self, req: Request) -> SolverClass:
        s = SolverClass(req.case, precision=self.precision)
        s.ReadRestartSolution(req.case.initial_conservative())
        if getattr(req.case, "sponge_elems", 0):
            s.SetDampingState(req.case.damping_primitive())
        return s

    def run(self, requests, clock=time.perf_counter) -> list:
        reqs = sorted(requests, key=lambda r: r.arrival)
        results = []
        queue = collections.deque()
        building = {}
        active = {}
        used = 0
        nxt = 0
        batch = Batch(tiles=self.tiles)
        pool = cf.ThreadPoolExecutor(max_workers=self.workers)
        t0 = clock()
        try:
            while nxt < len(reqs) or queue or building or active:
                now = clock() - t0
                while nxt < len(reqs) and reqs[nxt].arrival <= now:
                    queue.append(reqs[nxt])
                    nxt += 1
                # admission in arrival order by footprint
                while queue:
                    r = queue[0]
                    est = estimate_bytes(r.case)
                    if est > self.capacity:
                        queue.popleft()
                        results.append(Result(r.rid, r.kind, "rejected", r.arrival, finished=now,
                                              message=f"footprint {est} B exceeds the capacity"))
                        continue
                    if used + est > self.capacity:
                        break
                    queue.popleft()
                    used += est
                    building[pool.submit(self._build, r)] = (r, est, now)
                # built cases join the running batch
                for fut in [f for f in building if f.done()]:
                    r, est, t_adm = building.pop(fut)
                    try:
                        s = fut.result()
                    except Exception as exc:            # set-up rejected the case
                        used -= est
                        results.append(Result(r.rid, r.kind, "failed", r.arrival, admitted=t_adm,
                                              finished=clock() - t0, message=str(exc)))
                        continue
                    nbytes = s.device_bytes()
                    used += nbytes - est
                    batch.add(s, r.rk_steps)
                    active[s.h.value] = (r, s, nbytes, t_adm, clock() - t0)
                self.peak_bytes = max(self.peak_bytes, used)
                self.peak_active = max(self.peak_active, len(active))
                if not active:
                    if building:
                        cf.wait(list(building), timeout=0.05, return_when=cf.FIRST_COMPLETED)
                    elif nxt < len(reqs):
                        time.sleep(max(0.0, min(0.05, reqs[nxt].arrival - (clock() - t0))))
                    continue
                batch.step(self.steps_per_tick)
                status = batch.status()               # waits for the tick
                self.ticks += 1
                now = clock() - t0
                for s, words in status.items():
                    if words[ACTIVE]:
                        continue
                    r, s_, nbytes, t_adm, t_start = active.pop(s.h.value)
                    ok = words[FLAGS] == 0 and words[STEPS_DONE] == r.rk_steps
                    keep = self.keep_solutions
                    sol = s_.solution() if (keep is True or (keep and r.rid in keep)) else None
                    batch.remove(s_)
                    s_.close()
                    used -= nbytes
                    results.append(Result(r.rid, r.kind, "done" if ok else "failed", r.arrival, admitted=t_adm,
                                          started=t_start, finished=now, rk_steps=int(words[STEPS_DONE]),
                                          dofs=r.dofs, flags=int(words[FLAGS]), device_bytes=nbytes,
                                          message="" if ok else f"error flags {int(words[FLAGS])}",
                                          solution=sol))
        finally:
            pool.shutdown(wait=True)
            for _, s_, _, _, _ in active.values():
                s_.close()
            batch.close()
        results.sort(key=lambda x: x.rid)
        return results


def summary(results, makespan: float | None = None) -> dict:
    """p50/p99 latency of the completed requests and the aggregate rate."""
    done = [r for r in results if r.status == "done"]
    lat = np.array([r.latency for r in done]) if done else np.array([math.nan])
    if makespan is None:
        makespan = max(r.finished for r in done) - min(r.arrival for r in results) if done else math.nan
    gdof = sum(r.dofs * 3 * r.rk_steps for r in done) / 1e9
    return {"requests": len(results), "done": len(done),
            "failed": sum(r.status == "failed" for r in results),
            "rejected": sum(r.status == "rejected" for r in results),
            "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
            "mean_latency_s": float(lat.mean()), "makespan_s": float(makespan),
            "gdof_updates": gdof, "gdof_per_s": gdof / makespan if makespan else math.nan}


def serve_workload(seed: int = 4, count: int = 40, rate: float = 50.0, rk_steps: int = 200,
                   frac_tgv: float = 0.7, tgv_n=(4, 24), box_n=(48, 24, 24)) -> list:
    """Requests arriving as a Poisson stream (rate per second) from one seeded
    generator: with probability frac_tgv a Taylor-Green case of the batched
    workload's distribution (P uniform in 2..6, n^3 elements with n uniform
    in tgv_n, Mach U[0.05, 0.4], Re log-U[400, 3200], random phases), else
    the open box of box_n elements, P = 4, with PC-NSCBC inflow/outflow, an
    8-element sponge + CML at the outlet (fewer on short boxes), Mach
    U[0.3, 0.7] and vortex/pulse centres from the same generator."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for rid in range(count):
        t += float(rng.exponential(1.0 / rate))
        if rng.uniform() < frac_tgv:
            p = int(rng.integers(2, 7))
            n = int(rng.integers(tgv_n[0], tgv_n[1] + 1))
            case = TGVCase(n=n, p=p, rule=BATCH_RULE[p], mach=float(rng.uniform(0.05, 0.4)),
                           reynolds=float(math.exp(rng.uniform(math.log(400.0), math.log(3200.0)))),
                           phase=tuple(float(x) for x in rng.uniform(0.0, 2.0 * math.pi, 3)), name=f"tgv{rid}")
            kind = "tgv"
        else:
            sponge = min(8, max(1, box_n[0] // 6))
            case = OpenBoxCase(n=box_n, p=4, rule=12, mach=float(rng.uniform(0.3, 0.7)), sponge_elems=sponge,
                               seed=int(rng.integers(0, 2 ** 31)), name=f"box{rid}")
            kind = "openbox"
        out.append(Request(rid, case, rk_steps, t, kind))
    return out
