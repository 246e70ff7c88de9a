"""Flow integrals of a case's state for accuracy reports (host side, numpy).

The reference prints only the maximum Mach number and wall forces while it
runs (SolverClass::WriteTimeStepData, SolverClass.cpp:6174-6227); comparing
the mixed-precision form with FP64 needs integral quantities of the flow, so
this module evaluates them from the modal entropy variables with the case's
own Gauss rule, basis and metric terms:

    kinetic energy  KE = (1/|Omega|) int 1/2 rho |u|^2 dOmega
    enstrophy        Z = (1/|Omega|) int 1/2 rho |curl u|^2 dOmega

in the reference's non-dimensional units (rho, u, p scaled by rhoRef, uRef,
pRef; lengths as in the grid). The basis is the orthonormal Legendre basis
sqrt((2q+1)/2) P_q, DOF index i + K (j + K k) (ElementClass::mSol); the metric
terms (J and dr/dx .. dt/dz at the integration points) come from the
library's host set-up (vcp_host_setup), the same arrays the GPU path uses.
"""
from __future__ import annotations

import numpy as np

GAMMA = 1.4


def basis_tables(K: int, M: int):
    """Gauss points and weights (M) and the orthonormal Legendre values and
    derivatives at them, A[p][q] = phi_q(r_p), D[p][q] = phi_q'(r_p)."""
    r, w = np.polynomial.legendre.leggauss(M)
    A = np.zeros((M, K))
    D = np.zeros((M, K))
    for q in range(K):
        c = np.zeros(q + 1)
        c[q] = np.sqrt(0.5 * (2 * q + 1))
        A[:, q] = np.polynomial.legendre.legval(r, c)
        D[:, q] = np.polynomial.legendre.legval(r, np.polynomial.legendre.legder(c))
    return r, w, A, D


def _to_points(U, A, D):
    """Values and parametric (r, s, t) derivatives at the points of modal
    arrays U[e][v][k][j][i]; outputs [e][v][pk][pj][pi]."""
    val = np.einsum("ai,bj,ck,evkji->evcba", A, A, A, U, optimize=True)
    dr = np.einsum("ai,bj,ck,evkji->evcba", D, A, A, U, optimize=True)
    ds = np.einsum("ai,bj,ck,evkji->evcba", A, D, A, U, optimize=True)
    dt = np.einsum("ai,bj,ck,evkji->evcba", A, A, D, U, optimize=True)
    return val, dr, ds, dt


def flow_integrals(case, modal: np.ndarray, vol: np.ndarray | None = None, chunk: int = 1024) -> dict:
    """Kinetic energy, enstrophy (volume averages) and the domain volume of the
    modal entropy-variable state modal[nElem][5][(P+1)^3]. vol: the volume
    metric of vcp_host_setup ([nElem][13][M^3]: J, dr/dx .. dt/dz, x, y, z);
    computed here when not given."""
    K, M = case.p + 1, case.rule // 2 + 1
    if vol is None:
        from .solver import host_setup
        vol = host_setup(case)["vol"]
    _, w, A, D = basis_tables(K, M)
    w3 = np.einsum("c,b,a->cba", w, w, w)                 # [pk][pj][pi]
    E = modal.shape[0]
    ke = ens = volume = 0.0
    for e0 in range(0, E, chunk):
        e1 = min(E, e0 + chunk)
        U = modal[e0:e1].reshape(e1 - e0, 5, K, K, K)
        V, Vr, Vs, Vt = _to_points(U, A, D)
        g = vol[e0:e1].reshape(e1 - e0, 13, M, M, M)
        J = g[:, 0]
        c = g[:, 1:10]                                     # [r|s|t][x|y|z]
        grad = [Vr * c[:, 3 * 0 + d][:, None] + Vs * c[:, 3 * 1 + d][:, None] + Vt * c[:, 3 * 2 + d][:, None]
                for d in range(3)]                         # d/dx, d/dy, d/dz of V
        V4inv = 1.0 / V[:, 4]
        u = [-V[:, 1 + i] * V4inv for i in range(3)]
        ekin = 0.5 * (u[0] ** 2 + u[1] ** 2 + u[2] ** 2)
        s = GAMMA - (GAMMA - 1.0) * (V[:, 0] - V[:, 4] * ekin)
        rho = (-V[:, 4] * np.exp(s)) ** (-1.0 / (GAMMA - 1.0))
        # du_i/dx_d = -(1/V4) (dV_{1+i}/dx_d + u_i dV4/dx_d)
        du = [[-V4inv * (grad[d][:, 1 + i] + u[i] * grad[d][:, 4]) for d in range(3)] for i in range(3)]
        wx = du[2][1] - du[1][2]
        wy = du[0][2] - du[2][0]
        wz = du[1][0] - du[0][1]
        wJ = w3[None] * J
        ke += float(np.sum(wJ * rho * ekin))
        ens += float(np.sum(wJ * 0.5 * rho * (wx ** 2 + wy ** 2 + wz ** 2)))
        volume += float(np.sum(wJ))
    return dict(kinetic_energy=ke / volume, enstrophy=ens / volume, volume=volume)
