"""Python mirror of the reference SolverClass hot path over the C ABI
(include/vcp3d_amd.h), loading the in-tree gfx950 library.

Method names follow SolverClass (VCP3D_CurviLinear/include/SolverClass.hpp:
40-124): Residual, ExecuteTimeSynchronizationStep, ReadRestartSolution,
WriteRestartSolution. There is no CPU fallback: if the library is missing
the import fails loudly.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvcp3d_amd.so")


class CaseDesc(C.Structure):
    """vcp_case_desc (include/vcp3d_amd.h)."""
    _fields_ = [
        ("n_elem", C.c_int * 3),
        ("n_poly_grid", C.c_int),
        ("n_poly_sol", C.c_int),
        ("n_poly_int_rule", C.c_int),
        ("dof_location", C.c_int),
        ("periodic", C.c_int * 3),
        ("mach", C.c_double), ("reynolds", C.c_double),
        ("reynolds_length", C.c_double), ("temp_free_stream", C.c_double),
        ("theta_sym", C.c_double), ("rel_penalty", C.c_double),
        ("cfl", C.c_double), ("dt_synchr", C.c_double),
        ("bug_compat", C.c_int),
        ("bc", C.c_int * 6),
        ("bc_data", (C.c_double * 5) * 6),
        ("outlet_len", C.c_double), ("outlet_sigma", C.c_double),
        ("outlet_beta_l", C.c_double), ("outlet_beta_t", C.c_double),
        ("outlet_one_minus_beta", C.c_int), ("outlet_average_local", C.c_int),
        ("inlet_len", C.c_double), ("inlet_sigma", C.c_double),
        ("inlet_incoming", C.c_int),
        ("sponge_layer", C.c_int), ("cml", C.c_int),
        ("sponge_elems", C.c_int * 6),
        ("damping_constant", C.c_double * 6), ("damping_exponent", C.c_double * 6),
        ("monitor_forces", C.c_int), ("reference_area", C.c_double),
        ("compute_average", C.c_int), ("free_stream_dir", C.c_double * 3),
        ("sgs_model", C.c_int), ("body_force", C.c_double * 3), ("body_force_set", C.c_int),
    ]


EXPORTS = ["vcp_create", "vcp_destroy", "vcp_sizes", "vcp_reference", "vcp_set_restart",
           "vcp_set_damping_state", "vcp_nscbc_average",
           "vcp_set_solution", "vcp_get_solution", "vcp_get_nodal", "vcp_residual",
           "vcp_get_residual", "vcp_sync_step", "vcp_set_graph", "vcp_get_averages", "vcp_set_averages",
           "vcp_time_steps", "vcp_get_geometry", "vcp_cg_iterations",
           "vcp_profile_residual", "vcp_host_setup", "vcp_set_tile", "vcp_volume_tiles",
           "vcp_batch_create", "vcp_batch_destroy", "vcp_batch_add", "vcp_batch_remove",
           "vcp_batch_step", "vcp_batch_wait", "vcp_batch_status", "vcp_batch_set_graph",
           "vcp_batch_set_tile", "vcp_batch_groups", "vcp_batch_profile", "vcp_residual_static", "vcp_set_precision", "vcp_device_bytes",
           "vcp_damping_from_averages", "vcp_get_force_averages", "vcp_set_force_averages",
           "vcp_set_kernel_timing", "vcp_get_kernel_timing",
           "vcp_last_error"]

_lib = None


def load_library(path: str = LIB_PATH):
    """Loads libvcp3d_amd.so and declares the C ABI signatures."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(f"{path} is missing: build it with `python vcp3d_open_amd/build.py`")
    L = C.CDLL(path)
    dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
    ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
    vp = C.c_void_p
    L.vcp_create.argtypes = [C.POINTER(CaseDesc), dp, C.POINTER(vp)]
    L.vcp_destroy.argtypes = [vp]
    L.vcp_sizes.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.vcp_reference.argtypes = [vp, dp]
    L.vcp_set_restart.argtypes = [vp, dp]
    L.vcp_set_damping_state.argtypes = [vp, dp]
    L.vcp_nscbc_average.argtypes = [vp, dp]
    L.vcp_set_solution.argtypes = [vp, dp]
    L.vcp_get_solution.argtypes = [vp, dp]
    L.vcp_get_nodal.argtypes = [vp, dp]
    L.vcp_residual.argtypes = [vp, C.c_int, dp]
    L.vcp_get_residual.argtypes = [vp, dp]
    L.vcp_sync_step.argtypes = [vp, dp]
    L.vcp_set_graph.argtypes = [vp, C.c_int]
    L.vcp_get_averages.argtypes = [vp, dp, C.POINTER(C.c_int)]
    L.vcp_set_averages.argtypes = [vp, dp, C.c_int]
    L.vcp_time_steps.argtypes = [vp, dp]
    L.vcp_get_geometry.argtypes = [vp, dp, dp, dp]
    L.vcp_cg_iterations.argtypes = [vp, ip]
    L.vcp_profile_residual.argtypes = [vp, C.c_int, dp]
    L.vcp_host_setup.argtypes = [C.POINTER(CaseDesc), dp, dp, dp, dp, ip]
    L.vcp_set_tile.argtypes = [vp, C.c_int]
    L.vcp_volume_tiles.argtypes = [C.c_int, C.c_int, ip]
    L.vcp_batch_create.argtypes = [C.POINTER(vp)]
    L.vcp_batch_destroy.argtypes = [vp]
    L.vcp_batch_add.argtypes = [vp, vp, C.c_int]
    L.vcp_batch_remove.argtypes = [vp, vp]
    L.vcp_batch_step.argtypes = [vp, C.c_int]
    L.vcp_batch_wait.argtypes = [vp]
    L.vcp_batch_status.argtypes = [vp, C.POINTER(vp), ip, C.c_int]
    L.vcp_batch_set_graph.argtypes = [vp, C.c_int]
    L.vcp_batch_set_tile.argtypes = [vp, C.c_int, C.c_int, C.c_int]
    L.vcp_batch_groups.argtypes = [vp, ip, C.c_int]
    L.vcp_batch_profile.argtypes = [vp, C.c_int, dp, C.c_int]
    L.vcp_residual_static.argtypes = [vp, C.c_int, dp]
    L.vcp_set_precision.argtypes = [vp, C.c_int]
    L.vcp_device_bytes.argtypes = [vp, C.POINTER(C.c_size_t)]
    L.vcp_damping_from_averages.argtypes = [vp]
    L.vcp_get_force_averages.argtypes = [vp, dp]
    L.vcp_set_force_averages.argtypes = [vp, dp]
    L.vcp_set_kernel_timing.argtypes = [vp, C.c_int]
    L.vcp_get_kernel_timing.argtypes = [vp, dp, C.POINTER(C.c_int)]
    L.vcp_last_error.restype = C.c_char_p
    for name in EXPORTS:
        if name != "vcp_last_error":
            getattr(L, name).restype = C.c_int
    _lib = L
    return L


class VCPError(RuntimeError):
    pass


def desc_from_case(case, bug_compat: bool = True) -> CaseDesc:
    """vcp_case_desc from a case's description() dict (fields not named there
    stay 0)."""
    d = CaseDesc()
    for name, _ in CaseDesc._fields_:
        val = case.description().get(name) if name != "bug_compat" else None
        if val is None:
            continue
        attr = getattr(d, name)
        if hasattr(attr, "__len__"):
            for i, v in enumerate(val):
                if hasattr(attr[i], "__len__"):
                    for j, w in enumerate(v):
                        attr[i][j] = w
                else:
                    attr[i] = v
        else:
            setattr(d, name, val)
    d.bug_compat = 1 if bug_compat else 0
    return d


def host_setup(case, bug_compat: bool = True) -> dict:
    """Host-only tables/metrics of the library (no GPU needed)."""
    L = load_library()
    desc = desc_from_case(case, bug_compat)
    grid = np.ascontiguousarray(case.element_grid(), dtype=np.float64)
    E, M = case.n_elems, case.rule // 2 + 1
    vol = np.zeros((E, 13, M ** 3))
    faces = np.zeros((6, E, 16, M * M))
    lens = np.zeros((E, 4))
    nbr = np.zeros((E, 6), dtype=np.int32)
    rc = L.vcp_host_setup(C.byref(desc), grid, vol, faces, lens, nbr)
    if rc < 0:
        raise VCPError(L.vcp_last_error().decode())
    return dict(vol=vol, faces=faces, lens=lens, nbr=nbr)


class SolverClass:
    """One case on the GPU. Mirrors the reference SolverClass hot-path API."""

    def __init__(self, case, bug_compat: bool = True, elem_grid: np.ndarray | None = None,
                 use_graph: bool = False, precision: str = "fp64"):
        self._L = load_library()
        self.case = case
        desc = desc_from_case(case, bug_compat)
        grid = np.ascontiguousarray(case.element_grid() if elem_grid is None else elem_grid,
                                    dtype=np.float64)
        h = C.c_void_p()
        self._check(self._L.vcp_create(C.byref(desc), grid, C.byref(h)))
        self.h = h
        ne, k, m = C.c_int(), C.c_int(), C.c_int()
        self._L.vcp_sizes(h, C.byref(ne), C.byref(k), C.byref(m))
        self.n_elem, self.K, self.M = ne.value, k.value, m.value
        self.n_dofs = self.K ** 3
        if use_graph:
            self.set_graph(True)
        if precision != "fp64":
            self.set_precision(precision)

    def _check(self, rc: int) -> int:
        if rc < 0:
            raise VCPError(self._L.vcp_last_error().decode())
        return rc

    def close(self):
        if getattr(self, "h", None):
            self._L.vcp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- reference-named API ------------------------------------------------
    def ReadRestartSolution(self, nodal: np.ndarray) -> None:
        self._check(self._L.vcp_set_restart(self.h, np.ascontiguousarray(nodal, dtype=np.float64)))

    def Residual(self, compute_monitoring: bool = True) -> np.ndarray:
        md = np.zeros(8)
        self._check(self._L.vcp_residual(self.h, 1 if compute_monitoring else 0, md))
        return md

    def ExecuteTimeSynchronizationStep(self) -> np.ndarray:
        """Monitoring data; the number of RK steps taken is in last_rk_steps."""
        md = np.zeros(8)
        self.last_rk_steps = self._check(self._L.vcp_sync_step(self.h, md))
        return md

    def WriteRestartSolution(self) -> np.ndarray:
        """Nodal dimensional conservative payload of the restart file."""
        out = np.zeros((self.n_elem, self.n_dofs, 5))
        self._check(self._L.vcp_get_nodal(self.h, out))
        return out

    def SetDampingState(self, nodal_prim: np.ndarray) -> None:
        """Averaged nodal primitive state of the sponge layer
        (SolverClass::ReadAveragedSolutionSpongeLayer)."""
        self._check(self._L.vcp_set_damping_state(self.h, np.ascontiguousarray(nodal_prim, dtype=np.float64)))

    def DampingStateFromAverages(self) -> None:
        """Sponge target from the case's own running averages, on the device
        (in place of ReadAveragedSolutionSpongeLayer's averaged file)."""
        self._check(self._L.vcp_damping_from_averages(self.h))

    def AverageDataNSCBC(self) -> np.ndarray:
        """Boundary-averaged face-normal Mach number of the six block faces."""
        out = np.zeros(6)
        self._check(self._L.vcp_nscbc_average(self.h, out))
        return out

    # -- direct access ------------------------------------------------------
    def reference(self) -> dict:
        out = np.zeros(4)
        self._check(self._L.vcp_reference(self.h, out))
        return dict(u_ref=out[0], p_ref=out[1], rho_ref=out[2], mu=out[3])

    def averages(self):
        """Running averages [nElem][nDOFs][12] (the averaged columns of the
        reference's restart file) and the number of averaged steps."""
        out = np.zeros((self.n_elem, self.n_dofs, 12))
        n = C.c_int()
        self._check(self._L.vcp_get_averages(self.h, out, C.byref(n)))
        return out, n.value

    def set_averages(self, avg: np.ndarray, n_steps: int) -> None:
        self._check(self._L.vcp_set_averages(self.h, np.ascontiguousarray(avg, dtype=np.float64), int(n_steps)))

    def force_averages(self) -> np.ndarray:
        """Time-averaged force coefficients (mCfxAve, mCfyAve, mCfzAve,
        mCfxVisAve, mCfyVisAve, mCfzVisAve; SolverClass.cpp:3858-3869)."""
        out = np.zeros(6)
        self._check(self._L.vcp_get_force_averages(self.h, out))
        return out

    def set_force_averages(self, cf) -> None:
        """Continue the averaged force coefficients from restart meta data
        (SolverClass.cpp:3236-3249)."""
        cf = np.ascontiguousarray(cf, dtype=np.float64)
        if cf.shape != (6,):
            raise ValueError("six averaged force coefficients expected")
        self._check(self._L.vcp_set_force_averages(self.h, cf))

    def set_graph(self, enable: bool) -> None:
        """Replay each RK step of ExecuteTimeSynchronizationStep as a captured
        HIP graph (identical results; fewer launches)."""
        self._check(self._L.vcp_set_graph(self.h, 1 if enable else 0))

    def set_solution(self, modal: np.ndarray) -> None:
        self._check(self._L.vcp_set_solution(self.h, np.ascontiguousarray(modal, dtype=np.float64)))

    def solution(self) -> np.ndarray:
        out = np.zeros((self.n_elem, 5, self.n_dofs))
        self._check(self._L.vcp_get_solution(self.h, out))
        return out

    def residual_array(self) -> np.ndarray:
        out = np.zeros((self.n_elem, 5, self.n_dofs))
        self._check(self._L.vcp_get_residual(self.h, out))
        return out

    def time_steps(self) -> np.ndarray:
        out = np.zeros(self.n_elem)
        self._check(self._L.vcp_time_steps(self.h, out))
        return out

    def device_geometry(self) -> dict:
        """The metric terms computed on the device at creation: vgeo
        [nElem][10][M^3], fgeo [nElem][3][22][M^2], econst [nElem][16]."""
        M = self.M
        vg = np.zeros((self.n_elem, 10, M ** 3))
        fg = np.zeros((self.n_elem, 3, 22, M * M))
        ec = np.zeros((self.n_elem, 16))
        self._check(self._L.vcp_get_geometry(self.h, vg, fg, ec))
        return dict(vgeo=vg, fgeo=fg, econst=ec)

    def cg_iterations(self) -> np.ndarray:
        out = np.zeros(self.n_elem, dtype=np.int32)
        self._check(self._L.vcp_cg_iterations(self.h, out))
        return out

    def profile_residual(self, reps: int = 10) -> tuple:
        ms = np.zeros(2)
        self._check(self._L.vcp_profile_residual(self.h, reps, ms))
        return float(ms[0]), float(ms[1])

    def set_kernel_timing(self, enable: bool) -> None:
        """Time the face and volume launches of ExecuteTimeSynchronizationStep
        with HIP events on the library stream (enable clears the sums)."""
        self._check(self._L.vcp_set_kernel_timing(self.h, 1 if enable else 0))

    def kernel_timing(self) -> tuple:
        """(mean ms per face launch, mean ms per volume launch, timed stages)."""
        ms = np.zeros(2)
        n = C.c_int()
        self._check(self._L.vcp_get_kernel_timing(self.h, ms, C.byref(n)))
        return float(ms[0]), float(ms[1]), int(n.value)

    def device_bytes(self) -> int:
        """Device memory held by the context (bytes of its allocations)."""
        n = C.c_size_t()
        self._check(self._L.vcp_device_bytes(self.h, C.byref(n)))
        return int(n.value)

    def set_precision(self, precision: str) -> None:
        """"fp64" (the reference's arithmetic) or "mixed" (FP32 contractions
        and mass-matrix solve; state, pointwise physics, metrics, boundary
        kernels, time step and RK update in FP64)."""
        modes = {"fp64": 0, "mixed": 1}
        if precision not in modes:
            raise ValueError(f"precision must be one of {sorted(modes)}")
        self._check(self._L.vcp_set_precision(self.h, modes[precision]))
        self.precision = precision

    def set_tile(self, nt: int) -> None:
        """Volume workgroup size (threads per element; 0: the default)."""
        self._check(self._L.vcp_set_tile(self.h, int(nt)))


def volume_tiles(K: int, M: int) -> list:
    """Volume workgroup sizes available for (K, M), the default first."""
    L = load_library()
    out = np.zeros(4, dtype=np.int32)
    n = L.vcp_volume_tiles(K, M, out)
    if n < 0:
        raise VCPError(L.vcp_last_error().decode())
    return [int(x) for x in out[:n]]


# status words of vcp_batch_status
ACTIVE, STEPS_DONE, STEPS_LEFT, FLAGS, SYNC_DONE = range(5)


class Batch:
    """Many independent cases stepped together (vcp_batch_*): one launch per
    kernel and (K, M) group per RK stage, per-case time steps and
    synchronization-step clipping on the device. Members are SolverClass
    objects; a member's state is read with its own accessors after wait()."""

    def __init__(self, use_graph: bool = False, tiles: dict | None = None):
        self._L = load_library()
        h = C.c_void_p()
        self._check(self._L.vcp_batch_create(C.byref(h)))
        self.h = h
        self.members = []
        self.tiles = dict(tiles or {})
        if use_graph:
            self._check(self._L.vcp_batch_set_graph(self.h, 1))

    def _check(self, rc: int) -> int:
        if rc < 0:
            raise VCPError(self._L.vcp_last_error().decode())
        return rc

    def close(self):
        if getattr(self, "h", None):
            self._L.vcp_batch_destroy(self.h)
            self.h = None
            self.members = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, solver: "SolverClass", n_rk_steps: int = 0) -> None:
        self._check(self._L.vcp_batch_add(self.h, solver.h, int(n_rk_steps)))
        self.members.append(solver)
        nt = self.tiles.get((solver.K, solver.M))
        if nt:
            self._check(self._L.vcp_batch_set_tile(self.h, solver.K, solver.M, int(nt)))

    def remove(self, solver: "SolverClass") -> None:
        self._check(self._L.vcp_batch_remove(self.h, solver.h))
        self.members = [m for m in self.members if m is not solver]

    def set_tile(self, K: int, M: int, nt: int) -> None:
        self.tiles[(K, M)] = nt
        self._check(self._L.vcp_batch_set_tile(self.h, K, M, int(nt)))

    def step(self, n_steps: int = 1) -> None:
        """Queues n_steps RK steps of every active member (no wait)."""
        self._check(self._L.vcp_batch_step(self.h, int(n_steps)))

    def wait(self) -> None:
        self._check(self._L.vcp_batch_wait(self.h))

    def status(self) -> dict:
        """{solver: status words} (waits for the queued work)."""
        n = max(len(self.members), 1)
        ctxs = (C.c_void_p * n)()
        st = np.zeros((n, 5), dtype=np.int32)
        k = self._check(self._L.vcp_batch_status(self.h, ctxs, st, n))
        by_handle = {m.h.value: m for m in self.members}
        return {by_handle[ctxs[i]]: st[i].copy() for i in range(k)}

    def groups(self) -> list:
        info = np.zeros((64, 7), dtype=np.int32)
        n = self._check(self._L.vcp_batch_groups(self.h, info, 64))
        return [dict(K=int(r[0]), M=int(r[1]), bug=int(r[2]), cases=int(r[3]), work=int(r[4]),
                     nt=int(r[5]), mixed=int(r[6])) for r in info[:n]]

    def profile(self, reps: int = 5) -> list:
        """Per group: (ms of boundary + face kernels, ms of the volume kernel)."""
        ms = np.zeros(128)
        n = self._check(self._L.vcp_batch_profile(self.h, int(reps), ms, 64))
        return [(float(ms[2 * g]), float(ms[2 * g + 1])) for g in range(n)]
