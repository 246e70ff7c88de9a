"""Case builders: grids, initial conditions and parameter files.

The reference takes its grid from a Plot3D file and non-uniform initial states
from an SU2 restart (the SINE_AND_COSINE option aborts, ElementClass.cpp:526-532),
so every case here is expressed in those two inputs plus a parameter file.
Reference conditions follow InputParamClass::ComputeReferenceConditions
(VCP3D_CurviLinear/src/InputParamClass.cpp:605-758) for an all-periodic block:
mu = 1.716e-5 (constant), vInf = M sqrt(gamma R T), rhoRef = Re mu /(L vInf),
pRef = rhoRef R T, uRef = sqrt(pRef/rhoRef).
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field

import numpy as np

from . import io as vio

GAMMA = 1.4
R_GAS = 287.058
MU_DIM = 1.716e-5


def lgl_nodes(n: int) -> np.ndarray:
    """Legendre-Gauss-Lobatto nodes on [-1, 1] (n points), endpoints exact."""
    if n == 1:
        return np.zeros(1)
    if n == 2:
        return np.array([-1.0, 1.0])
    # interior nodes are the roots of P'_{n-1}
    c = np.zeros(n)
    c[-1] = 1.0
    dc = np.polynomial.legendre.legder(c)
    x = np.sort(np.polynomial.legendre.legroots(dc))
    # polish with Newton on P'_{n-1}
    d2c = np.polynomial.legendre.legder(dc)
    for _ in range(3):
        x = x - np.polynomial.legendre.legval(x, dc) / np.polynomial.legendre.legval(x, d2c)
    x = 0.5 * (x - x[::-1])  # enforce symmetry
    return np.concatenate([[-1.0], x, [1.0]])


def reference_conditions(mach: float, reynolds: float, length: float, temp: float):
    v_inf = mach * math.sqrt(GAMMA * R_GAS * temp)
    rho_ref = reynolds * MU_DIM / (length * v_inf)
    p_ref = rho_ref * R_GAS * temp
    u_ref = math.sqrt(p_ref / rho_ref)
    return dict(v_inf=v_inf, rho_ref=rho_ref, p_ref=p_ref, u_ref=u_ref,
                mu_nondim=MU_DIM * u_ref / p_ref)


# "Subgrid scale model" keyword of the parameter file and vcp_case_desc code
SGS_KEYWORD = {"none": "None", "wale": "WALE", "vreman": "Vreman"}
SGS_CODE = {"none": 0, "wale": 1, "vreman": 2}


def periodic_subfaces(n_elem, lengths):
    """Six all-periodic subface lines (translation = donor minus own face)."""
    ni, nj, nk = n_elem
    lx, ly, lz = lengths
    I, J, K = ni + 1, nj + 1, nk + 1
    fmt = lambda v: repr(float(v))
    return [
        [f"Periodic1to1Trans 1 1 1 {J} 1 {K}   {I} {I} 1 {J} 1 {K}   1 2 3   {fmt(lx)} 0.0 0.0"],
        [f"Periodic1to1Trans {I} {I} 1 {J} 1 {K}   1 1 1 {J} 1 {K}   1 2 3   {fmt(-lx)} 0.0 0.0"],
        [f"Periodic1to1Trans 1 {I} 1 1 1 {K}   1 {I} {J} {J} 1 {K}   1 2 3   0.0 {fmt(ly)} 0.0"],
        [f"Periodic1to1Trans 1 {I} {J} {J} 1 {K}   1 {I} 1 1 1 {K}   1 2 3   0.0 {fmt(-ly)} 0.0"],
        [f"Periodic1to1Trans 1 {I} 1 {J} 1 1   1 {I} 1 {J} {K} {K}   1 2 3   0.0 0.0 {fmt(lz)}"],
        [f"Periodic1to1Trans 1 {I} 1 {J} {K} {K}   1 {I} 1 {J} 1 1   1 2 3   0.0 0.0 {fmt(-lz)}"],
    ]


@dataclass
class TGVCase:
    """Taylor-Green vortex on the periodic cube [0, 2 pi]^3.

    u = V0 sin x cos y cos z, v = -V0 cos x sin y cos z, w = 0,
    p = p0 + rho0 V0^2/16 (cos 2x + cos 2y)(cos 2z + 2), rho = p/(R T0).
    With curvilinear=True the grid nodes are moved by
    x_i += 0.1 h sin(x) sin(y) sin(z) (evaluated at the unperturbed nodes),
    which leaves the periodic faces untouched.
    Optional phase shifts move the vortex: x -> x + phase[0], etc.
    """
    n: int = 8
    p: int = 3
    rule: int | None = None
    mach: float = 0.1
    reynolds: float = 1600.0
    temp: float = 300.0
    curvilinear: bool = False
    perturb: float = 0.1
    phase: tuple = (0.0, 0.0, 0.0)
    n_steps: int = 1
    dt_synchr: float | None = None
    cfl: float = 2.0
    theta_sym: float = 1.0
    rel_penalty: float = 1.0
    name: str = "tgv"
    sgs: str = "none"             # "Subgrid scale model": none, wale or vreman
    # "Body force per unit volume (x,y,z)", dimensional (None: keyword absent)
    body_force: tuple | None = None

    def __post_init__(self):
        if self.rule is None:
            self.rule = 3 * self.p if self.p <= 4 else {5: 15, 6: 16}.get(self.p, 3 * self.p)
        self.length = 2.0 * math.pi
        self.ref = reference_conditions(self.mach, self.reynolds, 1.0, self.temp)
        if self.dt_synchr is None:
            self.dt_synchr = 0.5 * self.stable_dt_estimate()

    @property
    def n_elem(self):
        return (self.n, self.n, self.n)

    @property
    def n_elems(self):
        return self.n ** 3

    @property
    def n_dofs(self):
        return (self.p + 1) ** 3

    def stable_dt_estimate(self) -> float:
        """Conservative dimensional time step for CFL = self.cfl, following the
        spectral-radius form of ElementClass::ComputeTimeStep
        (ElementClass.cpp:1522-1644) with bounds on |u| and a."""
        f1 = [2.0, 6.0, 12.0, 20.0, 28.0, 38.0, 50.0, 64.0, 80.0, 98.0]
        f2 = [6.0, 36.0, 150.0, 420.0, 980.0, 1975.0, 3575.0, 7000.0, 14000.0, 28000.0]
        h = self.length / self.n * (1.0 - 2.0 * self.perturb if self.curvilinear else 1.0)
        a = math.sqrt(GAMMA * R_GAS * self.temp) * 1.1
        v = self.ref["v_inf"] * 1.5
        rad_inv = 3.0 * (v + a) / h
        rho_min = 0.8 * self.ref["rho_ref"]
        rad_vis = MU_DIM / rho_min * max(2.0 - 2.0 / 3.0, GAMMA / 0.72) * 3.0 / h ** 2
        return self.cfl / (f1[self.p] * rad_inv + f2[self.p] * rad_vis)

    # -------------------------------------------------------------- geometry
    def grid_block(self) -> np.ndarray:
        """Plot3D block (3, kl, jl, il) with nodes at LGL positions."""
        r = lgl_nodes(self.p + 1)
        h = self.length / self.n
        pts = [0.0]
        for e in range(self.n):
            for q in range(1, self.p + 1):
                pts.append(e * h + 0.5 * (r[q] + 1.0) * h)
        x1 = np.array(pts)
        x1[-1] = self.length
        z, y, x = np.meshgrid(x1, x1, x1, indexing="ij")
        xyz = np.stack([x, y, z])
        if self.curvilinear:
            d = self.perturb * h * np.sin(x) * np.sin(y) * np.sin(z)
            xyz = xyz + d[None]
        return xyz

    def element_grid(self) -> np.ndarray:
        return vio.block_to_elements(self.grid_block(), self.n_elem, self.p)

    def nodal_coords(self) -> np.ndarray:
        """(nElem, nDOFs, 3): with equal grid and solution degree on LGL
        points the nodal solution DOFs coincide with the grid DOFs."""
        return np.transpose(self.element_grid(), (0, 2, 1)).copy()

    # ----------------------------------------------------- initial condition
    def initial_conservative(self, coords: np.ndarray | None = None) -> np.ndarray:
        """Dimensional conservative state (nElem, nDOFs, 5) at the nodes."""
        if coords is None:
            coords = self.nodal_coords()
        x = coords[..., 0] + self.phase[0]
        y = coords[..., 1] + self.phase[1]
        z = coords[..., 2] + self.phase[2]
        v0 = self.ref["v_inf"]
        rho0 = self.ref["rho_ref"]
        p0 = self.ref["p_ref"]
        u = v0 * np.sin(x) * np.cos(y) * np.cos(z)
        v = -v0 * np.cos(x) * np.sin(y) * np.cos(z)
        w = np.zeros_like(u)
        p = p0 + rho0 * v0 * v0 / 16.0 * (np.cos(2 * x) + np.cos(2 * y)) * (np.cos(2 * z) + 2.0)
        rho = p / (R_GAS * self.temp)
        rhoE = p / (GAMMA - 1.0) + 0.5 * rho * (u * u + v * v + w * w)
        return np.stack([rho, rho * u, rho * v, rho * w, rhoE], axis=-1)

    # ------------------------------------------------------ parameter file
    def params(self, grid_file="grid.xyz", restart_file="restart.dat",
               sol_file="out/Solution") -> dict:
        return {
            "Polynomial degree grid": self.p,
            "Polynomial degree solution": self.p,
            "Polynomial degree integration rule": self.rule,
            "Location of the DOFs": "LGL",
            "Grid file name": grid_file,
            "Solution file name": sol_file,
            "Restart file name": restart_file,
            "Restart": "yes",
            "Paraview visualization file name": "out/Solution.vtk",
            "Space average data file name": "out/SpaceAverage.vtk",
            "SU2 grid file name": "out/grid.su2",
            "Write SU2 grid file": "no",
            "Save solution every": 100000000,
            "Write time history info every": 1,
            "FEM working variables": "Entropy",
            "Riemann solver": "IsmailRoe",
            "Subgrid scale model": SGS_KEYWORD[self.sgs],
            "Wall modeling": "No wall model",
            "Theta parameter symmetrizing terms": self.theta_sym,
            "Relative penalty parameter": self.rel_penalty,
            "Mach number": self.mach,
            "Reynolds number": self.reynolds,
            "Reynolds length": 1.0,
            "Free stream temperature": self.temp,
            "Free stream velocity direction": [1.0, 0.0, 0.0],
            "Fluctuations to add to initial solution": "No fluctuations",
            "Monitor forces": "no",
            "Number of synchronization time steps": self.n_steps,
            "Synchronization time step": self.dt_synchr,
            "CFL number": self.cfl,
            "Compute average solution": "no",
            **({"Body force per unit volume (x,y,z)": [float(f) for f in self.body_force]}
               if self.body_force is not None else {}),
        }

    def subfaces(self):
        return periodic_subfaces(self.n_elem, (self.length,) * 3)

    def description(self) -> dict:
        """Fields of the case description (vcp_case_desc); omitted ones are 0."""
        return dict(n_elem=self.n_elem, n_poly_grid=self.p, n_poly_sol=self.p,
                    n_poly_int_rule=self.rule, dof_location=2, periodic=(1, 1, 1),
                    mach=self.mach, reynolds=self.reynolds, reynolds_length=1.0,
                    temp_free_stream=self.temp, theta_sym=self.theta_sym,
                    rel_penalty=self.rel_penalty, cfl=self.cfl, dt_synchr=self.dt_synchr,
                    sgs_model=SGS_CODE[self.sgs],
                    body_force=tuple(self.body_force) if self.body_force is not None else (0.0, 0.0, 0.0),
                    body_force_set=int(self.body_force is not None))

    def write_inputs(self, directory: str) -> str:
        """Write grid, restart and parameter file; returns the parameter path."""
        os.makedirs(os.path.join(directory, "out"), exist_ok=True)
        vio.write_plot3d(os.path.join(directory, "grid.xyz"), self.grid_block())
        vio.write_su2_restart(os.path.join(directory, "restart.dat"),
                              self.nodal_coords(), self.initial_conservative())
        path = os.path.join(directory, "param.cfg")
        vio.write_param_file(path, self.params(), self.subfaces())
        return path


# ------------------------------------------------------------------------
# Open box with characteristic in- and outflow
# ------------------------------------------------------------------------
BC_PERIODIC = 0
BC_OUTFLOW_NSCBC = 1
BC_INFLOW_STATIC_NSCBC = 2
BC_INFLOW_TOTAL_NSCBC = 3
_BC_KEYWORD = {BC_OUTFLOW_NSCBC: "BCOutflowNSCBC", BC_INFLOW_STATIC_NSCBC: "BCInflowStaticNSCBC",
               BC_INFLOW_TOTAL_NSCBC: "BCInflowTotalNSCBC"}


@dataclass
class OpenBoxCase:
    """Box [0, Lx] x [0, Ly] x [0, Lz] of uniform hexahedra, PC-NSCBC inflow
    at iMin and outflow at iMax, periodic in j and k, optional sponge layer
    with characteristic matching layer in front of the outlet.

    Free stream: Mach `mach` along +x at `temp`, `p_inf`. Initial state: an
    isentropic vortex (axis along z, strength eps = peak swirl speed / U,
    radius R) convected by the free stream plus a Gaussian pressure pulse
    p' = amp p_inf exp(-ln 2 |x - x_p|^2 / w^2) (rho' = p'/a^2), centres drawn
    from `seed`. The sponge target is the free stream.
    Reference conditions follow the prescribed boundary data: pRef = outlet
    pressure; rhoRef = inlet density, or pRef / (R T_t) for a total-condition
    inlet (InputParamClass.cpp:605-758).
    """
    n: tuple = (12, 6, 6)
    lengths: tuple = (2.0, 1.0, 1.0)
    p: int = 3
    rule: int | None = None
    mach: float = 0.5
    temp: float = 300.0
    p_inf: float = 101325.0
    inflow: str = "static"           # "static" or "total"
    sponge_elems: int = 2            # layers at the outlet (0: no sponge)
    cml: bool = True
    damping_constant: float = 20.0
    damping_exponent: float = 2.0
    outlet_len: float = 1.0
    outlet_sigma: float = 0.25
    outlet_beta_l: float = -1.0      # < 0: the boundary-averaged Mach number
    outlet_beta_t: float = -1.0
    outlet_one_minus_beta: bool = True
    inlet_len: float = 1.0
    inlet_sigma: float = 0.25
    inlet_incoming: bool = True
    vortex_strength: float = 0.5
    vortex_radius: float = 0.05      # fraction of Ly
    pulse_amp: float = 1.0e-3
    pulse_width: float = 0.05        # fraction of Ly
    seed: int = 0
    n_steps: int = 1
    dt_synchr: float | None = None
    cfl: float = 2.0
    theta_sym: float = 1.0
    rel_penalty: float = 1.0
    reynolds: float = 1.0e6          # written to the parameter file only
    compute_average: bool = False    # running averages of the primitive variables
    name: str = "openbox"
    sgs: str = "none"                # "Subgrid scale model": none, wale or vreman
    # "Body force per unit volume (x,y,z)", dimensional (None: keyword absent)
    body_force: tuple | None = None

    def __post_init__(self):
        if self.rule is None:
            self.rule = 3 * self.p if self.p <= 4 else {5: 15, 6: 16}.get(self.p, 3 * self.p)
        self.n = tuple(int(v) for v in self.n)
        a_inf = math.sqrt(GAMMA * R_GAS * self.temp)
        self.u_inf = self.mach * a_inf
        self.rho_inf = self.p_inf / (R_GAS * self.temp)
        f = 1.0 + 0.5 * (GAMMA - 1.0) * self.mach ** 2
        self.pt_inf = self.p_inf * f ** (GAMMA / (GAMMA - 1.0))
        self.tt_inf = self.temp * f
        rng = np.random.default_rng(self.seed)
        lx, ly, lz = self.lengths
        self.vortex_centre = (lx * rng.uniform(0.25, 0.45), ly * rng.uniform(0.4, 0.6))
        self.pulse_centre = (lx * rng.uniform(0.3, 0.6), ly * rng.uniform(0.3, 0.7), lz * rng.uniform(0.3, 0.7))
        p_ref = self.p_inf
        rho_ref = self.rho_inf if self.inflow == "static" else p_ref / (R_GAS * self.tt_inf)
        u_ref = math.sqrt(p_ref / rho_ref)
        self.ref = dict(p_ref=p_ref, rho_ref=rho_ref, u_ref=u_ref, mu_nondim=MU_DIM * u_ref / p_ref,
                        v_inf=self.u_inf)
        if self.dt_synchr is None:
            self.dt_synchr = 0.5 * self.stable_dt_estimate()

    @property
    def n_elem(self):
        return self.n

    @property
    def n_elems(self):
        return self.n[0] * self.n[1] * self.n[2]

    @property
    def n_dofs(self):
        return (self.p + 1) ** 3

    def stable_dt_estimate(self) -> float:
        f1 = [2.0, 6.0, 12.0, 20.0, 28.0, 38.0, 50.0, 64.0, 80.0, 98.0]
        h = min(L / n for L, n in zip(self.lengths, self.n))
        a = math.sqrt(GAMMA * R_GAS * self.temp) * 1.1
        v = self.u_inf * (1.0 + 1.5 * self.vortex_strength)
        rad_inv = 3.0 * (v + a) / h
        return self.cfl / (f1[self.p] * rad_inv)

    # -------------------------------------------------------------- geometry
    def grid_block(self) -> np.ndarray:
        r = lgl_nodes(self.p + 1)
        axes = []
        for L, n in zip(self.lengths, self.n):
            h = L / n
            pts = [0.0]
            for e in range(n):
                for q in range(1, self.p + 1):
                    pts.append(e * h + 0.5 * (r[q] + 1.0) * h)
            x1 = np.array(pts)
            x1[-1] = L
            axes.append(x1)
        z, y, x = np.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
        return np.stack([x, y, z])

    def element_grid(self) -> np.ndarray:
        return vio.block_to_elements(self.grid_block(), self.n_elem, self.p)

    def nodal_coords(self) -> np.ndarray:
        return np.transpose(self.element_grid(), (0, 2, 1)).copy()

    # ----------------------------------------------------- initial condition
    def initial_primitive(self, coords: np.ndarray | None = None) -> np.ndarray:
        """Dimensional (rho, u, v, w, p) at the nodes, (nElem, nDOFs, 5)."""
        if coords is None:
            coords = self.nodal_coords()
        lx, ly, lz = self.lengths
        x, y, z = coords[..., 0], coords[..., 1], coords[..., 2]
        xc, yc = self.vortex_centre
        R = self.vortex_radius * ly
        dx = x - xc
        dy = (y - yc + 0.5 * ly) % ly - 0.5 * ly          # nearest periodic image
        r2 = (dx * dx + dy * dy) / (R * R)
        f = np.exp(0.5 * (1.0 - r2))
        eps = self.vortex_strength
        u = self.u_inf * (1.0 - eps * dy / R * f)
        v = self.u_inf * eps * dx / R * f
        w = np.zeros_like(u)
        tr = 1.0 - 0.5 * (GAMMA - 1.0) * (self.mach * eps) ** 2 * f * f
        rho = self.rho_inf * tr ** (1.0 / (GAMMA - 1.0))
        p = self.p_inf * tr ** (GAMMA / (GAMMA - 1.0))
        xp, yp, zp = self.pulse_centre
        wd = self.pulse_width * ly
        d2 = (x - xp) ** 2 + ((y - yp + 0.5 * ly) % ly - 0.5 * ly) ** 2 + ((z - zp + 0.5 * lz) % lz - 0.5 * lz) ** 2
        dp = self.pulse_amp * self.p_inf * np.exp(-math.log(2.0) * d2 / (wd * wd))
        a2 = GAMMA * self.p_inf / self.rho_inf
        p = p + dp
        rho = rho + dp / a2
        return np.stack([rho, u, v, w, p], axis=-1)

    def initial_conservative(self, coords: np.ndarray | None = None) -> np.ndarray:
        q = self.initial_primitive(coords)
        rho, u, v, w, p = (q[..., i] for i in range(5))
        rhoE = p / (GAMMA - 1.0) + 0.5 * rho * (u * u + v * v + w * w)
        return np.stack([rho, rho * u, rho * v, rho * w, rhoE], axis=-1)

    def damping_primitive(self) -> np.ndarray:
        """Sponge target (the averaged primitive state): the free stream."""
        q = np.zeros((self.n_elems, self.n_dofs, 5))
        q[..., 0] = self.rho_inf
        q[..., 1] = self.u_inf
        q[..., 4] = self.p_inf
        return q

    # ------------------------------------------------------ boundary data
    def bc(self) -> list:
        inflow = BC_INFLOW_STATIC_NSCBC if self.inflow == "static" else BC_INFLOW_TOTAL_NSCBC
        return [inflow, BC_OUTFLOW_NSCBC, 0, 0, 0, 0]

    def bc_data(self) -> list:
        data = [[0.0] * 5 for _ in range(6)]
        if self.inflow == "static":
            data[0][:4] = [self.rho_inf, self.u_inf, 0.0, 0.0]
        else:
            data[0][:] = [self.pt_inf, self.tt_inf, 1.0, 0.0, 0.0]
        data[1][0] = self.p_inf
        return data

    def description(self) -> dict:
        d = dict(n_elem=self.n_elem, n_poly_grid=self.p, n_poly_sol=self.p,
                 n_poly_int_rule=self.rule, dof_location=2, periodic=(0, 1, 1),
                 mach=self.mach, reynolds=self.reynolds, reynolds_length=1.0,
                 temp_free_stream=self.temp, theta_sym=self.theta_sym,
                 rel_penalty=self.rel_penalty, cfl=self.cfl, dt_synchr=self.dt_synchr,
                 bc=self.bc(), bc_data=self.bc_data(),
                 outlet_len=self.outlet_len, outlet_sigma=self.outlet_sigma,
                 outlet_beta_l=self.outlet_beta_l, outlet_beta_t=self.outlet_beta_t,
                 outlet_one_minus_beta=int(self.outlet_one_minus_beta), outlet_average_local=0,
                 inlet_len=self.inlet_len, inlet_sigma=self.inlet_sigma,
                 inlet_incoming=int(self.inlet_incoming), compute_average=int(self.compute_average),
                 sgs_model=SGS_CODE[self.sgs],
                 body_force=tuple(self.body_force) if self.body_force is not None else (0.0, 0.0, 0.0),
                 body_force_set=int(self.body_force is not None))
        if self.sponge_elems > 0:
            d.update(sponge_layer=1, cml=int(self.cml),
                     sponge_elems=[0, self.sponge_elems, 0, 0, 0, 0],
                     damping_constant=[self.damping_constant] * 6,
                     damping_exponent=[self.damping_exponent] * 6)
        return d

    # ------------------------------------------------------ parameter file
    def params(self, grid_file="grid.xyz", restart_file="restart.dat",
               sol_file="out/Solution") -> dict:
        prm = TGVCase.params(self, grid_file, restart_file, sol_file)
        prm.update({
            "Reynolds number": self.reynolds,
            "Outlet characteristic length": self.outlet_len,
            "Outlet normal relaxation (sigma)": self.outlet_sigma,
            "Outlet use one minus beta transverse relaxation": "yes" if self.outlet_one_minus_beta else "no",
            "Outlet use local (element) averaged Mach number": "no",
            "Outlet coupled   transverse relaxation (beta_l)": self.outlet_beta_l,
            "Outlet uncoupled transverse relaxation (beta_t)": self.outlet_beta_t,
            "Inlet characteristic length": self.inlet_len,
            "Inlet normal relaxation (sigma)": self.inlet_sigma,
            "Inlet use outgoing wave amplitude": "yes" if self.inlet_incoming else "no",
        })
        if self.sponge_elems > 0:
            prm["Averaged solution file name"] = "averaged.dat"
            prm["Sponge layer"] = "yes"
            prm["Characteristic matching layer"] = "yes" if self.cml else "no"
        return prm

    def subfaces(self):
        ni, nj, nk = self.n
        lx, ly, lz = self.lengths
        I, J, K = ni + 1, nj + 1, nk + 1
        fmt = lambda v: repr(float(v))
        bc = self.bc()
        return [
            [f"{_BC_KEYWORD[bc[0]]} 1 1 1 {J} 1 {K}"],
            [f"{_BC_KEYWORD[bc[1]]} {I} {I} 1 {J} 1 {K}"],
            [f"Periodic1to1Trans 1 {I} 1 1 1 {K}   1 {I} {J} {J} 1 {K}   1 2 3   0.0 {fmt(ly)} 0.0"],
            [f"Periodic1to1Trans 1 {I} {J} {J} 1 {K}   1 {I} 1 1 1 {K}   1 2 3   0.0 {fmt(-ly)} 0.0"],
            [f"Periodic1to1Trans 1 {I} 1 {J} 1 1   1 {I} 1 {J} {K} {K}   1 2 3   0.0 0.0 {fmt(lz)}"],
            [f"Periodic1to1Trans 1 {I} 1 {J} {K} {K}   1 {I} 1 {J} 1 1   1 2 3   0.0 0.0 {fmt(-lz)}"],
        ]

    def prescribed(self) -> list:
        if self.inflow == "static":
            inlet = ["1 1 1 Density VelocityX VelocityY VelocityZ",
                     f"    {self.rho_inf!r} {self.u_inf!r} 0.0 0.0"]
        else:
            inlet = ["1 1 1 TotalPressure TotalTemperature VelocityDirX VelocityDirY VelocityDirZ",
                     f"    {self.pt_inf!r} {self.tt_inf!r} 1.0 0.0 0.0"]
        return inlet + ["2 1 1 Pressure", f"    {self.p_inf!r}"]

    def sponge_block(self) -> str | None:
        if self.sponge_elems <= 0:
            return None
        lines = ["Sponge layer information"]
        for f in range(6):
            nb = self.sponge_elems if f == 1 else 0
            lines.append(f"{f + 1} {nb} {self.damping_constant!r} {self.damping_exponent!r}")
        return "\n".join(lines)

    def write_inputs(self, directory: str) -> str:
        os.makedirs(os.path.join(directory, "out"), exist_ok=True)
        coords = self.nodal_coords()
        vio.write_plot3d(os.path.join(directory, "grid.xyz"), self.grid_block())
        vio.write_su2_restart(os.path.join(directory, "restart.dat"), coords, self.initial_conservative(coords))
        if self.sponge_elems > 0:
            q = self.damping_primitive()
            cons = np.zeros_like(q)
            extra = np.concatenate([q, np.zeros(q.shape[:2] + (7,))], axis=2)
            vio.write_su2_restart(os.path.join(directory, "averaged.dat"), coords, cons, extra=extra,
                                  extra_names=["AvDensity", "AvVelocity_x", "AvVelocity_y", "AvVelocity_z",
                                               "AvPressure", "AvVxVx", "AvVyVy", "AvVzVz", "AvVxVy",
                                               "AvVxVz", "AvVyVz", "AvEddyVis"])
        path = os.path.join(directory, "param.cfg")
        vio.write_param_file(path, self.params(), self.subfaces(), self.prescribed(), self.sponge_block())
        return path


# ------------------------------------------------------------------------
# Batched workload: many independent Taylor-Green cases
# ------------------------------------------------------------------------
BATCH_RULE = {2: 6, 3: 9, 4: 12, 5: 15, 6: 16}


def batch_workload(seed: int = 2024, count: int = 64, n_range=(4, 24), p_range=(2, 6),
                   curvilinear: bool = False) -> list:
    """Taylor-Green cases with per-case values drawn from one seeded generator:
    P uniform in p_range, rule 3P (P <= 4), 15 (P = 5), 16 (P = 6); n^3
    elements with n uniform in n_range; Mach uniform in [0.05, 0.4]; Reynolds
    log-uniform in [400, 3200]; phase shifts uniform in [0, 2 pi)^3."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        p = int(rng.integers(p_range[0], p_range[1] + 1))
        n = int(rng.integers(n_range[0], n_range[1] + 1))
        mach = float(rng.uniform(0.05, 0.4))
        re = float(math.exp(rng.uniform(math.log(400.0), math.log(3200.0))))
        phase = tuple(float(x) for x in rng.uniform(0.0, 2.0 * math.pi, 3))
        out.append(TGVCase(n=n, p=p, rule=BATCH_RULE[p], mach=mach, reynolds=re, phase=phase,
                           curvilinear=curvilinear, name=f"tgv{i}"))
    return out
