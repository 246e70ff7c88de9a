"""File formats shared with VCP3D: Plot3D grids, SU2 binary restarts and the
keyword parameter file.

Formats follow the reference readers/writers:
  * Plot3D multi-block Fortran-unformatted grid, one block:
    VCP3D_CurviLinear/src/SolverClass.cpp:2787-3034 (ReadGrid) and
    VCP3D_CurviLinear/include/SolverClass.hpp:646-667 (ReadGridCoordinates).
  * SU2 binary restart: int[5] header {535532, nVar, nDOFsTot, 1, 8}, nVar
    names of 33 chars, per element (k-major, then j, then i) per nodal DOF
    [x, y, z, rho, rho u, rho v, rho w, rho E], then the time-step int and 8
    metadata doubles: SolverClass.cpp:4375-4804 (WriteRestartSolution) and
    SolverClass.cpp:3040-3376 (ReadRestartSolution).
  * Parameter file: case-insensitive "keyword: value" lines plus the
    "Subface information" block: InputParamClass.cpp:155-440, :1441-1576.
"""
from __future__ import annotations

import struct
import numpy as np

SU2_MAGIC = 535532
NAME_LEN = 33


# ----------------------------------------------------------------- Plot3D
def write_plot3d(path: str, xyz: np.ndarray) -> None:
    """Write one block of shape (3, kl, jl, il) (i fastest on disk)."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    _, kl, jl, il = xyz.shape
    nbytes = 8 * 3 * il * jl * kl
    with open(path, "wb") as f:
        f.write(struct.pack("<3i", 4, 1, 4))
        f.write(struct.pack("<5i", 12, il, jl, kl, 12))
        f.write(struct.pack("<i", nbytes))
        f.write(xyz.tobytes(order="C"))
        f.write(struct.pack("<i", nbytes))


def read_plot3d(path: str) -> np.ndarray:
    """Read a single-block Plot3D file written as above (double precision)."""
    with open(path, "rb") as f:
        data = f.read()
    rec = struct.unpack_from("<i", data, 0)[0]
    off = 4
    if rec == 4:  # multi-block header
        nblk, _ = struct.unpack_from("<2i", data, off)
        if nblk != 1:
            raise ValueError("only single-block Plot3D grids are supported")
        off += 8
        off += 4  # leading int of the block-size record
    il, jl, kl = struct.unpack_from("<3i", data, off)
    off += 12 + 4
    nbytes = struct.unpack_from("<i", data, off)[0]
    off += 4
    npts = il * jl * kl
    prec = nbytes // (3 * npts)
    dt = np.float64 if prec == 8 else np.float32
    arr = np.frombuffer(data, dtype=dt, count=3 * npts, offset=off)
    return arr.astype(np.float64).reshape(3, kl, jl, il)


def block_to_elements(xyz: np.ndarray, n_elem, n_poly_grid: int) -> np.ndarray:
    """Split a (3, kl, jl, il) block into per-element grid DOFs.

    Returns (nElem, 3, nGridDOFs) with the element index i fastest, then j,
    then k, and the grid DOF index ii fastest, then jj, then kk, exactly as
    SolverClass::ReadGrid fills mCoorNodalGridDOFs (SolverClass.cpp:2988-3034).
    """
    ni, nj, nk = n_elem
    p = n_poly_grid
    out = np.empty((nk, nj, ni, 3, p + 1, p + 1, p + 1))
    for k in range(nk):
        for j in range(nj):
            for i in range(ni):
                blk = xyz[:, k * p:k * p + p + 1, j * p:j * p + p + 1, i * p:i * p + p + 1]
                out[k, j, i] = blk
    return out.reshape(ni * nj * nk, 3, (p + 1) ** 3)


# ------------------------------------------------------------ SU2 restart
def write_su2_restart(path: str, coords: np.ndarray, cons: np.ndarray,
                      time_step: int = 0, extra: np.ndarray | None = None,
                      extra_names: list | None = None,
                      meta: np.ndarray | None = None) -> None:
    """Write a restart file.

    coords: (nElem, nDOFs, 3) nodal coordinates; cons: (nElem, nDOFs, 5)
    dimensional conservative variables; element order i fastest, then j, k.
    extra: optional further columns (nElem, nDOFs, n), e.g. the averaged
    primitive variables, velocity products and eddy viscosity of an averaged
    solution file (SolverClass.hpp:748-809 reads columns 8..12).
    meta: the 8 meta data doubles after the time step, the first six the
    time-averaged force coefficients (SolverClass.cpp:4407-4409); zeros when
    not given.
    """
    n_elem, n_dofs, _ = cons.shape
    names = ["x", "y", "z", "Density", "Momentum_x", "Momentum_y",
             "Momentum_z", "Energy"]
    cols = [coords, cons]
    if extra is not None:
        cols.append(extra)
        names += list(extra_names) if extra_names else [f"Extra_{i}" for i in range(extra.shape[2])]
    buf = np.concatenate(cols, axis=2).astype(np.float64)
    with open(path, "wb") as f:
        f.write(struct.pack("<5i", SU2_MAGIC, len(names), n_elem * n_dofs, 1, 8))
        for nm in names:
            f.write(nm.encode().ljust(NAME_LEN, b"\0"))
        f.write(buf.tobytes(order="C"))
        f.write(struct.pack("<i", time_step))
        md = np.zeros(8)
        if meta is not None:
            meta = np.asarray(meta, dtype=np.float64).ravel()
            if meta.size > 8:
                raise ValueError("at most 8 restart meta data doubles")
            md[:meta.size] = meta
        f.write(md.tobytes())


def read_su2_restart(path: str, n_dofs: int, with_meta: bool = False):
    """Read a restart file; returns (names, data[nElem, nDOFs, nVar], step),
    and the 8 meta data doubles as a fourth item with with_meta (zeros when
    the file ends after the time step)."""
    with open(path, "rb") as f:
        data = f.read()
    hdr = struct.unpack_from("<5i", data, 0)
    if hdr[0] != SU2_MAGIC:
        raise ValueError("not an SU2 restart file")
    nvar, ndof_tot = hdr[1], hdr[2]
    off = 20
    names = []
    for _ in range(nvar):
        names.append(data[off:off + NAME_LEN].split(b"\0")[0].decode())
        off += NAME_LEN
    nbytes = len(data) - off - 4
    prec = round(nbytes / (nvar * ndof_tot + 8))
    dt = np.float64 if prec == 8 else np.float32
    arr = np.frombuffer(data, dtype=dt, count=nvar * ndof_tot, offset=off)
    arr = arr.astype(np.float64).reshape(ndof_tot // n_dofs, n_dofs, nvar)
    off += nvar * ndof_tot * prec
    step = struct.unpack_from("<i", data, off)[0]
    if not with_meta:
        return names, arr, step
    off += 4
    meta = np.zeros(8)
    n_meta = min(8, (len(data) - off) // prec)
    if n_meta > 0:
        meta[:n_meta] = np.frombuffer(data, dtype=dt, count=n_meta, offset=off)
    return names, arr, step, meta


# --------------------------------------------------------- parameter file
def write_param_file(path: str, params: dict, subfaces: list,
                     prescribed: list | None = None,
                     sponge: str | None = None) -> None:
    """Write a parameter file.

    params: keyword -> value (written "keyword: value").
    subfaces: six lists (one per block face, iMin..kMax) of subface lines.
    prescribed: optional lines of the "Prescribed boundary data" block.
    sponge: optional raw text of the "Sponge layer information" block.
    """
    lines = []
    for key, val in params.items():
        if isinstance(val, (list, tuple)):
            val = " ".join(repr(float(v)) if isinstance(v, float) else str(v) for v in val)
        elif isinstance(val, float):
            val = repr(val)
        lines.append(f"{key}: {val}")
    if sponge:
        lines.append("")
        lines.append(sponge)
    lines.append("")
    lines.append("Subface information")
    for face, sfs in enumerate(subfaces):
        lines.append(f"{face + 1} {len(sfs)}")
        lines.extend("    " + s for s in sfs)
    if prescribed:
        lines.append("")
        lines.append("Prescribed boundary data")
        lines.extend(prescribed)
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
