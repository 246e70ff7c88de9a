"""A VCP3D case deck as the reference reads it: the parameter file, the Plot3D
grid, the SU2 restart (or the free-stream start plus fluctuations) and, for a
sponge layer, the averaged-solution file.

    case = DeckCase("param.cfg")
    solver = open_solver("param.cfg")       # create + initial state + sponge target
    for _ in range(case.n_sync_steps):
        solver.ExecuteTimeSynchronizationStep()
    write_restart(solver, case, "Solution_TimeStep_100", 100)

Parsing follows InputParamClass:
  * keyword lines: text after '#' dropped, "keyword: value" split at the first
    ':', blanks trimmed, keyword compared in lower case
    (InputParamClass.cpp:155-183); a later line overrides an earlier one;
  * "Subface information": per block face a line "<face> <nSubfaces>" (lines
    whose first token is not that face number are skipped), then one line per
    subface "<BC> ranges ..." (InputParamClass.cpp:1441-1576);
  * "Prescribed boundary data": "<face> <subface> <nPoints> <names...>" then
    nPoints value lines (InputParamClass.cpp:1340-1398,
    SubfaceClasses.cpp:304-350);
  * "Sponge layer information": per face "<face> <nElem> <sigma> <alpha>"
    (InputParamClass.cpp:1196-1330), read when a sponge layer or CML is on.
Defaults are the InputParamClass constructor's (InputParamClass.cpp:12-121).

The initial state follows the reference main (VCP3D_CurviLinear.cpp:50-90):
InitSol sets the free stream (InputParamClass::DetermineNonDimPrimVarFreeStream,
InputParamClass.cpp:991-1191: from the prescribed in/outflow data when they fix
the state, else from Mach number, temperature, direction and Reynolds number),
ReadRestartSolution replaces it when "Restart: yes", and AddFluctuations adds
the pressure pulse while the time step is 0 (ElementClass.cpp:515-665).

Only what the GPU path supports is accepted (entropy variables, Ismail-Roe, no
wall model; block faces periodic, PC-NSCBC in/outflow,
subsonic total-condition inflow, subsonic outflow, supersonic inflow and
outflow, heat-flux, isothermal or inviscid walls, farfield, one subface each);
anything else raises DeckError naming the setting.
"""
from __future__ import annotations

import math
import os

import numpy as np

from . import io as vio
from .cases import GAMMA, MU_DIM, R_GAS

# "Subgrid scale model" (InputParamClass.cpp:241-250) -> vcp_case_desc.sgs_model
SGS_CODES = {"none": 0, "wale": 1, "vreman": 2}
BC_CODES = {"periodic1to1trans": 0, "bcoutflownscbc": 1, "bcinflowstaticnscbc": 2,
            "bcinflowtotalnscbc": 3, "bcinflowsubsonic": 4, "bcoutflowsubsonic": 5,
            "bcwallheatflux": 6, "bcwallisothermal": 7, "bcwallinviscid": 8,
            "bcinflowsupersonic": 9, "bcoutflowsupersonic": 10, "bcfarfield": 11}
# prescribed data per boundary condition, in the order of vcp_case_desc.bc_data
_BC_DATA = {1: ["Pressure"], 2: ["Density", "VelocityX", "VelocityY", "VelocityZ"],
            3: ["TotalPressure", "TotalTemperature", "VelocityDirX", "VelocityDirY", "VelocityDirZ"],
            4: ["TotalPressure", "TotalTemperature", "VelocityDirX", "VelocityDirY", "VelocityDirZ"],
            5: ["Pressure"], 6: ["HeatFlux"], 7: ["Temperature"], 8: [],
            9: ["Density", "VelocityX", "VelocityY", "VelocityZ", "Pressure"], 10: [], 11: []}
_KNOWN_BC = ("internal1to1", "periodic1to1trans", "bcfarfield", "bcwallisothermal",
             "bcwallisothermalwm", "bcwallheatflux", "bcwallheatfluxwm", "bcwallinviscid",
             "bcsymmetryplane", "bcinflowsubsonic", "bcinflowsupersonic", "bcoutflowsubsonic",
             "bcoutflowsupersonic", "bcinflowstaticnscbc", "bcinflowtotalnscbc", "bcoutflownscbc")
_VAL_LARGE = 1.0e300


class DeckError(ValueError):
    """A parameter file the GPU path cannot run, with the setting named."""


def _blank(line: str) -> str:
    return line.replace("\t", " ").replace("\r", " ").strip()


class ParamFile:
    """The parameter file: keyword values, subfaces, prescribed data, sponge data."""

    def __init__(self, path: str):
        self.path = os.path.abspath(path)
        with open(path) as f:
            self.lines = f.read().splitlines()
        self.values = {}
        for raw in self.lines:
            line = raw.replace("\t", " ").replace("\r", " ")
            if "#" in line:
                line = line[:line.index("#")]
            if ":" not in line:
                continue
            key, val = line.split(":", 1)
            key, val = key.strip().lower(), val.strip()
            if key and val:
                self.values[key] = val
        self.subfaces = self._read_subfaces()
        self.prescribed = self._read_prescribed()
        self.sponge = None
        if self.yes("sponge layer", False) or self.yes("characteristic matching layer", False):
            self.sponge = self._read_sponge()

    # ------------------------------------------------------------ values
    def has(self, key: str) -> bool:
        return key in self.values

    def text(self, key: str, default: str | None = None) -> str:
        if key in self.values:
            return self.values[key]
        if default is None:
            raise DeckError(f'"{key}" not specified in {self.path}')
        return default

    def number(self, key: str, default: float | None = None) -> float:
        if key not in self.values:
            if default is None:
                raise DeckError(f'"{key}" not specified in {self.path}')
            return default
        return float(self.values[key].split()[0])

    def integer(self, key: str, default: int | None = None) -> int:
        return int(self.number(key, None if default is None else float(default)))

    def vector(self, key: str, default=None) -> list:
        if key not in self.values:
            return list(default) if default is not None else None
        return [float(t) for t in self.values[key].split()[:3]]

    def yes(self, key: str, default: bool) -> bool:
        if key not in self.values:
            return default
        v = self.values[key].lower()
        if v not in ("yes", "no"):
            raise DeckError(f'"{key}" must be yes or no, not "{self.values[key]}"')
        return v == "yes"

    # ------------------------------------------------------------ blocks
    def _block_start(self, title: str) -> int | None:
        for i, raw in enumerate(self.lines):
            if _blank(raw).lower() == title:
                return i + 1
        return None

    def _read_subfaces(self) -> list:
        i = self._block_start("subface information")
        if i is None:
            raise DeckError(f'keyword "Subface information" not found in {self.path}')
        faces = []
        for face in range(6):
            n = None
            while i < len(self.lines):
                tok = self.lines[i].split()
                i += 1
                if tok and tok[0].lstrip("+-").isdigit() and int(tok[0]) == face + 1:
                    n = int(tok[1])
                    break
            if n is None:
                raise DeckError(f"subface information for face {face + 1} not found")
            subs = []
            for _ in range(n):
                tok = self.lines[i].split()
                i += 1
                if not tok or tok[0].lower() not in _KNOWN_BC:
                    raise DeckError(f'unknown boundary condition "{tok[0] if tok else ""}" on face {face + 1}')
                subs.append(dict(bc=tok[0].lower(), name=tok[0], ints=[int(t) for t in tok[1:7]],
                                 rest=tok[7:]))
            faces.append(subs)
        return faces

    def _read_prescribed(self) -> dict:
        i = self._block_start("prescribed boundary data")
        out = {}
        if i is None:
            return out
        while i < len(self.lines):
            tok = self.lines[i].split()
            i += 1
            try:
                face, sub, npts = int(tok[0]), int(tok[1]), int(tok[2])
            except (ValueError, IndexError):
                continue
            names = tok[3:]
            rows = []
            for _ in range(npts):
                rows.append([float(t) for t in self.lines[i].split()[:len(names)]])
                i += 1
            out[(face - 1, sub - 1)] = {nm: [r[k] for r in rows] for k, nm in enumerate(names)}
        return out

    def _read_sponge(self) -> list:
        i = self._block_start("sponge layer information")
        if i is None:
            raise DeckError(f'keyword "Sponge layer information" not found in {self.path}')
        rows = []
        for face in range(6):
            found = None
            while i < len(self.lines):
                tok = self.lines[i].split()
                i += 1
                if tok and tok[0].lstrip("+-").isdigit():
                    if int(tok[0]) != face + 1:
                        raise DeckError(f"sponge layer is not specified for face {face + 1}")
                    found = (int(tok[1]), float(tok[2]), float(tok[3]))
                    break
            if found is None:
                raise DeckError(f"sponge layer is not specified for face {face + 1}")
            if found[0] < 0:
                raise DeckError(f"negative sponge layer thickness on face {face + 1}")
            rows.append(found)
        return rows


class DeckCase:
    """A case built from a parameter file; offers what the solver and the
    oracle take from a case (description(), element_grid()) plus the initial
    state and the sponge target of the reference's set-up sequence."""

    def __init__(self, param_path: str):
        prm = ParamFile(param_path)
        self.prm = prm
        self.dir = os.path.dirname(prm.path)
        self._check_supported(prm)
        self.p = prm.integer("polynomial degree solution")
        self.p_grid = prm.integer("polynomial degree grid")
        self.rule = prm.integer("polynomial degree integration rule")
        loc = prm.text("location of the dofs").lower()
        if loc not in ("lgl", "equidistant"):
            raise DeckError('unknown "Location of the DOFs"')
        self.dof_location = 2 if loc == "lgl" else 1
        self.mach = prm.number("mach number", -1.0)
        self.reynolds = prm.number("reynolds number", -1.0)
        self.reynolds_length = prm.number("reynolds length", -1.0)
        self.temp = prm.number("free stream temperature", -1.0)
        self.vel_dir = prm.vector("free stream velocity direction")
        self.dt_synchr = prm.number("synchronization time step")
        self.cfl = prm.number("cfl number")
        self.n_sync_steps = prm.integer("number of synchronization time steps", -1)
        self.theta_sym = prm.number("theta parameter symmetrizing terms", 1.0)
        self.monitor_forces = prm.yes("monitor forces", True)
        self.compute_average = prm.yes("compute average solution", False)
        self.sgs_model = SGS_CODES[prm.text("subgrid scale model", "none").lower()]
        # "Body force per unit volume (x,y,z)" (InputParamClass.cpp:331-334):
        # the keyword's presence switches the source term on
        self.body_force = prm.vector("body force per unit volume (x,y,z)")
        # InputParamClass.cpp:599: no continuation without averaging
        self.continue_averaging = self.compute_average and prm.yes("continue averaging from restart", False)
        self.reference_area = prm.number("reference area force coefficients", 1.0)
        self.rel_penalty = prm.number("relative penalty parameter", 1.0)
        self.grid_file = os.path.join(self.dir, prm.text("grid file name"))
        self.restart = prm.yes("restart", False)
        self.restart_file = os.path.join(self.dir, prm.text("restart file name", "")) if self.restart else None
        self.sol_file = os.path.join(self.dir, prm.text("solution file name", "Solution.dat"))
        self._grid_block = vio.read_plot3d(self.grid_file)
        _, kl, jl, il = self._grid_block.shape
        pg = self.p_grid
        if any((n - 1) % pg for n in (il, jl, kl)):
            raise DeckError("grid point counts are not a multiple of the grid degree plus one")
        self.n = ((il - 1) // pg, (jl - 1) // pg, (kl - 1) // pg)
        self._boundaries(prm)
        self._sponge(prm)
        self._fluctuations(prm)

    # ------------------------------------------------------------ checks
    @staticmethod
    def _check_supported(prm: ParamFile):
        need = {"fem working variables": "entropy", "riemann solver": "ismailroe"}
        for key, val in need.items():
            if prm.text(key, "").lower() != val:
                raise DeckError(f'"{key}": only {val} is supported on this path')
        if prm.text("subgrid scale model", "none").lower() not in SGS_CODES:
            # InputParamClass.cpp:241-250
            raise DeckError('Unknown "Subgrid scale model" specified (None, WALE or Vreman)')
        if prm.text("wall modeling", "no wall model").lower() != "no wall model":
            raise DeckError('"Wall modeling": only "No wall model" is supported on this path')
        if prm.yes("outlet use local (element) averaged mach number", False):
            raise DeckError('"Outlet use local (element) averaged Mach number: yes" is not supported')

    def _boundaries(self, prm: ParamFile):
        self.bc = [0] * 6
        self.bc_data = [[0.0] * 5 for _ in range(6)]
        self.periodic = [1, 1, 1]
        for f, subs in enumerate(prm.subfaces):
            if len(subs) != 1:
                raise DeckError(f"face {f + 1}: one subface per block face is supported")
            bc = subs[0]["bc"]
            if bc not in BC_CODES:
                raise DeckError(f'face {f + 1}: boundary condition "{subs[0]["name"]}" is not supported')
            if bc == "periodic1to1trans":
                continue
            self.periodic[f // 2] = 0
            self.bc[f] = BC_CODES[bc]
            data = prm.prescribed.get((f, 0), {})
            if not _BC_DATA[self.bc[f]] and data:
                raise DeckError(f"face {f + 1}: boundary data is prescribed, but this is not expected")
            if self.bc[f] == 4 and "TotalPressure" not in data:
                raise DeckError(f"face {f + 1}: BCInflowSubsonic is supported with total conditions "
                                "(TotalPressure, TotalTemperature, VelocityDir*)")
            for q, nm in enumerate(_BC_DATA[self.bc[f]]):
                if nm not in data:
                    raise DeckError(f'face {f + 1}: prescribed "{nm}" missing')
                if len(set(data[nm])) != 1:
                    raise DeckError(f"face {f + 1}: spatially varying prescribed data is not supported")
                self.bc_data[f][q] = data[nm][0]
        for d in range(3):
            if self.bc[2 * d] == 0 and self.bc[2 * d + 1] == 0 and not self.periodic[d]:
                raise DeckError(f"direction {d}: unsupported pairing of block faces")
        self.nscbc = dict(
            outlet_len=prm.number("outlet characteristic length", 1.0),
            outlet_sigma=prm.number("outlet normal relaxation (sigma)", 0.0),
            outlet_beta_l=prm.number("outlet coupled   transverse relaxation (beta_l)", 1.0),
            outlet_beta_t=prm.number("outlet uncoupled transverse relaxation (beta_t)", 1.0),
            outlet_one_minus_beta=int(prm.yes("outlet use one minus beta transverse relaxation", True)),
            outlet_average_local=0,
            inlet_len=prm.number("inlet characteristic length", 1.0),
            inlet_sigma=prm.number("inlet normal relaxation (sigma)", 1.0),
            inlet_incoming=int(prm.yes("inlet use outgoing wave amplitude", True)))

    def _sponge(self, prm: ParamFile):
        self.sponge_layer = prm.yes("sponge layer", False)
        self.cml = prm.yes("characteristic matching layer", False)
        self.averaged_file = None
        if prm.sponge is None:
            self.sponge_rows = [(0, 0.0, 0.0)] * 6
        else:
            self.sponge_rows = prm.sponge
        if self.sponge_layer or self.cml:
            self.averaged_file = os.path.join(self.dir, prm.text("averaged solution file name"))
        self.sponge_elems = max(r[0] for r in self.sponge_rows) if (self.sponge_layer or self.cml) else 0

    def _fluctuations(self, prm: ParamFile):
        kind = prm.text("fluctuations to add to initial solution", "no fluctuations").lower()
        if kind not in ("no fluctuations", "pressure pulse"):
            raise DeckError(f'"Fluctuations to add to initial solution: {kind}" is not supported')
        self.pulse = None
        if kind == "pressure pulse":
            self.pulse = dict(centre=prm.vector("pressure pulse center", (0.0, 0.0, 0.0)),
                              strength=prm.number("pressure pulse strength", 0.5),
                              width=prm.number("pressure pulse width", 0.5),
                              one_d=prm.yes("one-dimensional pulse", False),
                              direction=self._pulse_direction(prm))

    @staticmethod
    def _pulse_direction(prm: ParamFile) -> float:
        """"Pulse (x-)direction": upstream -1, downstream +1 (default +1)."""
        v = prm.text("pulse (x-)direction", "downstream").lower()
        if v not in ("upstream", "downstream"):
            raise DeckError('unknown "Pulse (x-)direction" (upstream or downstream)')
        return -1.0 if v == "upstream" else 1.0

    # ------------------------------------------------------------ case API
    @property
    def n_elem(self):
        return self.n

    @property
    def n_elems(self):
        return self.n[0] * self.n[1] * self.n[2]

    @property
    def n_dofs(self):
        return (self.p + 1) ** 3

    def description(self) -> dict:
        d = dict(n_elem=self.n, n_poly_grid=self.p_grid, n_poly_sol=self.p,
                 n_poly_int_rule=self.rule, dof_location=self.dof_location,
                 periodic=tuple(self.periodic), mach=self.mach, reynolds=self.reynolds,
                 reynolds_length=self.reynolds_length, temp_free_stream=self.temp,
                 theta_sym=self.theta_sym, rel_penalty=self.rel_penalty, cfl=self.cfl,
                 dt_synchr=self.dt_synchr, bc=self.bc, bc_data=self.bc_data,
                 monitor_forces=int(self.monitor_forces), reference_area=self.reference_area,
                 compute_average=int(self.compute_average), sgs_model=self.sgs_model,
                 body_force=tuple(self.body_force) if self.body_force is not None else (0.0, 0.0, 0.0),
                 body_force_set=int(self.body_force is not None),
                 free_stream_dir=self.vel_dir if self.vel_dir is not None else (1.0, 0.0, 0.0), **self.nscbc)
        if self.sponge_layer or self.cml:
            d.update(sponge_layer=int(self.sponge_layer), cml=int(self.cml),
                     sponge_elems=[r[0] for r in self.sponge_rows],
                     damping_constant=[r[1] for r in self.sponge_rows],
                     damping_exponent=[r[2] for r in self.sponge_rows])
        return d

    def element_grid(self) -> np.ndarray:
        return vio.block_to_elements(self._grid_block, self.n, self.p_grid)

    def nodal_coords(self) -> np.ndarray:
        """(nElem, nDOFs, 3) coordinates of the nodal solution DOFs: the grid
        DOFs when grid and solution degree agree on LGL points."""
        if self.p_grid != self.p:
            raise DeckError("nodal coordinates need equal grid and solution degrees")
        return np.transpose(self.element_grid(), (0, 2, 1)).copy()

    def free_stream_primitive(self) -> np.ndarray:
        """Dimensional (rho, u, v, w, p) of the free stream
        (DetermineNonDimPrimVarFreeStream, before its scaling)."""
        vals = {}
        for data in self.prm.prescribed.values():
            for nm, v in data.items():
                vals[nm] = v[0]
        get = lambda nm: vals.get(nm, -1.0)
        p, p_tot, rho, rho_tot = get("Pressure"), get("TotalPressure"), get("Density"), get("TotalDensity")
        T, T_tot = get("Temperature"), get("TotalTemperature")
        vel = [vals.get(nm, _VAL_LARGE) for nm in ("VelocityX", "VelocityY", "VelocityZ")]
        vdir = [vals.get(nm, _VAL_LARGE) for nm in ("VelocityDirX", "VelocityDirY", "VelocityDirZ")]
        gm1 = GAMMA - 1.0
        if all(v < _VAL_LARGE for v in vel):
            n = sum(x > 0.0 for x in (p, rho, T))
            if n >= 2:
                if p <= 0.0:
                    p = rho * R_GAS * T
                if rho <= 0.0:
                    rho = p / (R_GAS * T)
                return np.array([rho, *vel, p])
            if n == 1 and T <= 0.0:
                vm2 = sum(v * v for v in vel)
                if p <= 0.0:
                    p = 0.95 * rho * vm2 / GAMMA
                else:
                    rho = GAMMA * p / (0.95 * vm2)
                return np.array([rho, *vel, p])
        elif all(v < _VAL_LARGE for v in vdir):
            inv = 1.0 / math.sqrt(sum(v * v for v in vdir))
            vdir = [v * inv for v in vdir]
            m2 = 0.95 * 0.95
            if p_tot > 0.0 and p > 0.0:
                m2 = ((p_tot / p) ** (gm1 / GAMMA) - 1.0) * 2.0 / gm1
            elif rho_tot > 0.0 and rho > 0.0:
                m2 = ((rho_tot / rho) ** (1.0 / GAMMA) - 1.0) * 2.0 / gm1
            elif T_tot > 0.0 and T > 0.0:
                m2 = (T_tot / T - 1.0) * 2.0 / gm1
            if sum(x > 0.0 for x in (p_tot, rho_tot, T_tot)) >= 2:
                if p_tot <= 0.0:
                    p_tot = rho_tot * R_GAS * T_tot
                if rho_tot <= 0.0:
                    rho_tot = p_tot / (R_GAS * T_tot)
                tmp = 1.0 + 0.5 * gm1 * m2
                p = p_tot / tmp ** (GAMMA / gm1)
                rho = rho_tot / tmp ** (1.0 / gm1)
                vmag = math.sqrt(m2 * GAMMA * p / rho)
                return np.array([rho, vdir[0] * vmag, vdir[1] * vmag, vdir[2] * vmag, p])
        # from the free-stream data of the parameter file
        if self.mach <= 0.0 or self.reynolds <= 0.0 or self.reynolds_length <= 0.0 or self.temp < 0.0:
            raise DeckError("free stream not determined: Mach number, Reynolds number, "
                            "Reynolds length and Free stream temperature are needed")
        if self.vel_dir is None:
            raise DeckError('"Free stream velocity direction" not specified')
        inv = 1.0 / math.sqrt(sum(v * v for v in self.vel_dir))
        v_inf = self.mach * math.sqrt(GAMMA * R_GAS * self.temp)
        rho_inf = self.reynolds * MU_DIM / (self.reynolds_length * v_inf)
        p_inf = rho_inf * R_GAS * self.temp
        return np.array([rho_inf, *(v_inf * d * inv for d in self.vel_dir), p_inf])

    def initial_primitive(self) -> np.ndarray:
        """Dimensional (rho, u, v, w, p) at the nodal DOFs, (nElem, nDOFs, 5):
        the restart or the free stream, plus the pressure pulse at time step 0."""
        step = 0
        if self.restart:
            _, data, step = vio.read_su2_restart(self.restart_file, self.n_dofs)
            cons = data[:, :, 3:8]
            rho = cons[..., 0]
            vel = cons[..., 1:4] / rho[..., None]
            p = (GAMMA - 1.0) * (cons[..., 4] - 0.5 * rho * (vel * vel).sum(axis=-1))
            q = np.concatenate([rho[..., None], vel, p[..., None]], axis=-1)
        else:
            q = np.broadcast_to(self.free_stream_primitive(), (self.n_elems, self.n_dofs, 5)).copy()
        if self.pulse is not None and step == 0:
            xyz = self.nodal_coords()
            c = self.pulse["centre"]
            d0, w0 = self.pulse["strength"], self.pulse["width"]
            with np.errstate(divide="ignore"):
                cc = np.float64(0.5) / np.float64(w0 * w0)   # inf for a zero width, as in the reference
            dx, dy, dz = (xyz[..., i] - c[i] for i in range(3))
            if self.pulse["one_d"]:
                pinf, rinf = q[..., 4].copy(), q[..., 0]
                gg = pinf * d0 * np.exp(-dx * dx * cc)
                q[..., 4] += gg
                q[..., 1] += self.pulse["direction"] * gg / np.sqrt(GAMMA * pinf * rinf)
            else:
                q[..., 4] = q[..., 4] * (1.0 + d0 * np.exp(-(dx * dx + dy * dy + dz * dz) * cc))
        return q

    def initial_conservative(self) -> np.ndarray:
        """Dimensional conservative state at the nodal DOFs, (nElem, nDOFs, 5):
        the restart payload as stored when no fluctuation is added to it."""
        if self.restart:
            _, data, step = vio.read_su2_restart(self.restart_file, self.n_dofs)
            if self.pulse is None or step != 0:
                return np.ascontiguousarray(data[:, :, 3:8])
        q = self.initial_primitive()
        rho, u, v, w, p = (q[..., i] for i in range(5))
        rhoE = p / (GAMMA - 1.0) + 0.5 * rho * (u * u + v * v + w * w)
        return np.stack([rho, rho * u, rho * v, rho * w, rhoE], axis=-1)

    def damping_primitive(self) -> np.ndarray:
        """Dimensional averaged primitive state (rho, u, v, w, p) at the nodal
        DOFs from the averaged-solution file: its columns 8..12
        (SolverClass::ReadAveragedSolutionSpongeLayer, SolverClass.hpp:748-809)."""
        if self.averaged_file is None:
            raise DeckError("no sponge layer or characteristic matching layer in this deck")
        _, data, _ = vio.read_su2_restart(self.averaged_file, self.n_dofs)
        if data.shape[2] < 13:
            raise DeckError("averaged solution file has fewer than 13 variables")
        return np.ascontiguousarray(data[:, :, 8:13])


def open_solver(param_path: str, bug_compat: bool = True, **kw):
    """The reference main's set-up on the GPU path: SolverClass from the deck,
    initial state (restart or free stream plus fluctuations) and, with a sponge
    layer, its averaged target state. Returns (solver, case)."""
    from .solver import SolverClass
    case = DeckCase(param_path)
    s = SolverClass(case, bug_compat=bug_compat, **kw)
    s.ReadRestartSolution(case.initial_conservative())
    if case.sponge_layer or case.cml:
        s.SetDampingState(case.damping_primitive())
    if case.continue_averaging and case.restart:
        # the averaged columns, the number of averaged steps and the averaged
        # force coefficients of the restart (SolverClass.cpp:3227-3249, :4401-4414)
        _, data, n, meta = vio.read_su2_restart(case.restart_file, case.n_dofs, with_meta=True)
        if data.shape[2] < 20:
            raise DeckError("restart file without averaged columns for \"Continue averaging from restart\"")
        s.set_averages(data[:, :, 8:20], n)
        s.set_force_averages(meta[:6])
    return s, case


AVERAGE_NAMES = ["Density_Average", "VelocityX_Average", "VelocityY_Average", "VelocityZ_Average",
                 "Pressure_Average", "VelocityX_VelocityX_Average", "VelocityY_VelocityY_Average",
                 "VelocityZ_VelocityZ_Average", "VelocityX_VelocityY_Average", "VelocityX_VelocityZ_Average",
                 "VelocityY_VelocityZ_Average", "EddyVis_Average"]


def write_restart(solver, case, path: str, time_step: int) -> None:
    """SolverClass::WriteRestartSolution (SolverClass.cpp:4375-4804): the
    nodal dimensional conservative state with the DOF coordinates as an SU2
    restart file the reference can read back; with running averages their
    twelve columns follow, the stored step count is the number of averaged
    steps and the meta data hold the averaged force coefficients, as the
    reference writes them (SolverClass.cpp:4401-4414)."""
    extra = names = meta = None
    if getattr(case, "compute_average", False):
        extra, time_step = solver.averages()
        names = AVERAGE_NAMES
        meta = solver.force_averages()
    vio.write_su2_restart(path, case.nodal_coords(), solver.WriteRestartSolution(), time_step=time_step,
                          extra=extra, extra_names=names, meta=meta)
