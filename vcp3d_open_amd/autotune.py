"""Volume workgroup-size autotuner for batched stepping.

The volume kernel of each (nDOFs1D, nInt1D) shape is compiled for several
workgroup sizes (threads per element; vcp_volume_tiles lists them). Which one
is fastest depends on the shape: a small element (K = 3) leaves most of a
256-thread workgroup idle in its pencil stages, while a large one (K = 7) needs
512 threads to hold one DOF per thread. The tuner times every size of every
group of a batch on the batch's own cases (vcp_batch_profile: residual
evaluations that store nothing, HIP events around each launch) and keeps the
fastest. Sizes give results equal to rounding (the mass-matrix solve groups
its partial sums differently), and a given size gives identical results
every run.

The choice is recorded per device architecture in tiles_<arch>.json beside
this file: {"K,M": {"nt": best, "ms": {nt: volume ms per launch}}}.
"""
from __future__ import annotations

import json
import os

from .solver import Batch, volume_tiles

HERE = os.path.dirname(os.path.abspath(__file__))


def table_path(arch: str = "gfx950") -> str:
    return os.path.join(HERE, f"tiles_{arch}.json")


def load_tiles(arch: str = "gfx950") -> dict:
    """{(K, M): nt} from the recorded table ({} when there is none)."""
    try:
        with open(table_path(arch)) as f:
            data = json.load(f)
    except (OSError, ValueError):
        return {}
    out = {}
    for key, rec in data.items():
        k, m = (int(x) for x in key.split(","))
        out[(k, m)] = int(rec["nt"])
    return out


def tune(batch: Batch, reps: int = 5) -> dict:
    """Times every volume workgroup size of every (K, M) shape of the batch
    (summed over the groups of that shape); sets the fastest on the batch and
    returns {(K, M): {"nt": best, "ms": {nt: volume ms per launch}}}."""
    shapes = []
    for g in batch.groups():
        if (g["K"], g["M"]) not in shapes:
            shapes.append((g["K"], g["M"]))
    result = {}
    for K, M in shapes:
        times = {}
        for nt in volume_tiles(K, M):
            batch.set_tile(K, M, nt)
            prof = batch.profile(reps)
            times[nt] = sum(ms_vol for g, (_, ms_vol) in zip(batch.groups(), prof)
                            if (g["K"], g["M"]) == (K, M))
        best = min(times, key=times.get)
        batch.set_tile(K, M, best)
        result[(K, M)] = {"nt": best, "ms": times}
    return result


def save(result: dict, arch: str = "gfx950") -> str:
    """Writes the tuner's result as the recorded table."""
    data = {f"{k},{m}": {"nt": rec["nt"], "ms": {str(nt): ms for nt, ms in rec["ms"].items()}}
            for (k, m), rec in sorted(result.items())}
    path = table_path(arch)
    with open(path, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    return path
