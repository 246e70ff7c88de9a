/*
 * vcp_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement (plain C) of the VCP3D explicit DG residual and SSP-RK3
 * synchronization step, used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker. The product (vcp3d_open_amd)
 * never links, loads or calls it.
 *
 * Scope: entropy working variables, Ismail-Roe flux, no SGS model or the
 * WALE / Vreman model, no wall model, structured single block whose faces are periodic (translation),
 * PC-NSCBC in/outflow, subsonic (total-condition) inflow, subsonic outflow,
 * or heat-flux / isothermal / inviscid walls, with optional sponge layer and
 * characteristic matching layer.
 * Each function cites the reference file:line it restates; paths are
 * relative to VCP3D_CurviLinear/.
 */
#ifndef VCP_ORACLE_H
#define VCP_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int    n_elem[3];        /* elements in i, j, k */
  int    n_poly_grid;      /* "Polynomial degree grid" */
  int    n_poly_sol;       /* "Polynomial degree solution" */
  int    n_poly_int_rule;  /* "Polynomial degree integration rule" */
  int    dof_location;     /* 1 equidistant, 2 LGL ("Location of the DOFs") */
  int    periodic[3];      /* 1: periodic translation in i/j/k */
  double mach, reynolds, reynolds_length, temp_free_stream;
  double theta_sym, rel_penalty, cfl, dt_synchr;
  int    bug_compat;       /* 1: metric indexed by variable index, as the
                              reference does (ElementClass.cpp:3865-3873,
                              FluxesFace.cpp:272-282, :325-332), and the
                              transverse derivatives of the reconstructed
                              NSCBC state stored in swapped slots
                              (BoundaryState.cpp:1261-1287) */
  /* Physical block faces, used where periodic[face/2] == 0; faces are
     iMin, iMax, jMin, jMax, kMin, kMax (one subface per block face). */
  int    bc[6];            /* ORC_BC_* */
  double bc_data[6][5];    /* dimensional prescribed data of the face:
                              NSCBC or subsonic outflow {p}; static NSCBC
                              inflow {rho, u, v, w}; total NSCBC or subsonic
                              inflow {p_t, T_t, dir_x, dir_y, dir_z};
                              heat-flux wall {q}; isothermal wall {T};
                              supersonic inflow {rho, u, v, w, p} */
  /* "NSCBC tuning parameters" (InputParamClass.cpp:406-424) */
  double outlet_len, outlet_sigma, outlet_beta_l, outlet_beta_t;
  int    outlet_one_minus_beta, outlet_average_local;
  double inlet_len, inlet_sigma;
  int    inlet_incoming;
  /* "Sponge layer", "Characteristic matching layer" and the "Sponge layer
     information" block (InputParamClass.cpp:1196-1330) */
  int    sponge_layer, cml;
  int    sponge_elems[6];
  double damping_constant[6], damping_exponent[6];
  /* "Monitor forces", "Reference area force coefficients" (<= 0: 1) */
  int    monitor_forces;
  double reference_area;
  int    compute_average;  /* "Compute average solution" */
  double free_stream_dir[3]; /* "Free stream velocity direction" */
  int    sgs_model;        /* "Subgrid scale model": ORC_SGS_* */
  /* "Body force per unit volume (x,y,z)", dimensional; body_force_set = 0:
     the keyword absent (InputParamClass.cpp:331-334) */
  double body_force[3];
  int    body_force_set;
} orc_params;

enum { ORC_SGS_NONE = 0, ORC_SGS_WALE = 1, ORC_SGS_VREMAN = 2 };

enum { ORC_BC_PERIODIC = 0, ORC_BC_OUTFLOW_NSCBC = 1,
       ORC_BC_INFLOW_STATIC_NSCBC = 2, ORC_BC_INFLOW_TOTAL_NSCBC = 3,
       ORC_BC_INFLOW_SUBSONIC = 4,      /* total conditions */
       ORC_BC_OUTFLOW_SUBSONIC = 5, ORC_BC_WALL_HEATFLUX = 6,
       ORC_BC_WALL_ISOTHERMAL = 7, ORC_BC_WALL_INVISCID = 8,
       ORC_BC_INFLOW_SUPERSONIC = 9, ORC_BC_OUTFLOW_SUPERSONIC = 10, ORC_BC_FARFIELD = 11 };

typedef struct orc_case orc_case;

/* elem_grid: [nElem][3][(n_poly_grid+1)^3], element i fastest, then j, k;
   grid DOF ii fastest (the reference's mCoorNodalGridDOFs). */
orc_case *orc_create(const orc_params *prm, const double *elem_grid);
void      orc_destroy(orc_case *c);
const char *orc_last_error(void);

/* Sizes: nElem, K (=P+1), M (integration points 1D). */
void orc_sizes(const orc_case *c, int *n_elem, int *n_dofs1d, int *n_int1d);
/* uRef, pRef, rhoRef, mu (non-dimensional) as ComputeReferenceConditions. */
void orc_reference(const orc_case *c, double out[4]);

/* nodal: [nElem][nDOFs][5] dimensional conservative (restart layout). */
int  orc_set_restart(orc_case *c, const double *nodal);
/* modal: [nElem][5][nDOFs] entropy variables. */
void orc_set_modal(orc_case *c, const double *modal);
void orc_get_modal(const orc_case *c, double *modal);
/* residual after the mass-matrix solve, [nElem][5][nDOFs]. */
void orc_get_residual(const orc_case *c, double *res);
/* nodal dimensional conservative [nElem][nDOFs][5] (restart writer). */
void orc_get_nodal(const orc_case *c, double *nodal);
/* running averages [nElem][nDOFs][12] (rho, u, v, w, p, uu, vv, ww, uv, uw,
   vw, eddy viscosity), dimensional, and the number of averaged steps */
int  orc_get_averages(const orc_case *c, double *avg, int *n_steps);

/* SolverClass::Residual; returns 0 or <0 on error. */
int  orc_residual(orc_case *c, int compute_monitoring, double monitoring[8]);
/* SolverClass::ExecuteTimeSynchronizationStep; returns number of RK steps. */
int  orc_sync_step(orc_case *c, double monitoring[8]);
/* Per-element stable time step (ElementClass::ComputeTimeStep). */
int  orc_time_steps(orc_case *c, double *dt_elem);
/* CG iteration count of the last mass-matrix solve, per element. */
void orc_cg_iterations(const orc_case *c, int *iters);

/* Sponge-layer target: the averaged nodal primitive state (rho, u, v, w, p),
   dimensional, [nElem][nDOFs][5] (SolverClass::ReadAveragedSolutionSpongeLayer,
   SpongeLayer.cpp:665-838). Required before the first residual when
   sponge_layer is set. */
int  orc_set_damping_state(orc_case *c, const double *nodal_prim);
/* SolverClass::AverageDataNSCBC: boundary-averaged Mach number of the six
   block faces for the current state (0 where no NSCBC face). */
void orc_nscbc_average(orc_case *c, double mach_avg[6]);

/* Setup data for tests. which: 0 volume metric [nElem][13][nInt],
   1..6 face metric (iMin,iMax,jMin,jMax,kMin,kMax) [nElem][16][nInt2D],
   7 length scales+volume [nElem][4] (I,J,K,vol), 8 neighbours [nElem][6] (as double),
   9 damping functions [nElem][3][nInt] (0 outside the sponge),
   10 damping state [nElem][5][nInt] (0 outside the sponge). */
void orc_get_setup(const orc_case *c, int which, double *out);
/* Test hook: replace the volume metric (0), a face metric (1..6) or the
   length scales (7) with given arrays of the orc_get_setup layouts, so that
   the residual can be evaluated on another set-up's geometry. */
void orc_set_setup(orc_case *c, int which, const double *in);

/* Eddy viscosity rho nu_t of subgrid scale model `model` (ORC_SGS_*) for the
   LES length scale len, density rho and velocity gradients
   g = (du/dx, dv/dx, dw/dx, du/dy, dv/dy, dw/dy, du/dz, dv/dz, dw/dz)
   (SGSModel.cpp:34-210). */
double orc_eddy_viscosity(int model, double len, double rho, const double g[9]);
/* 1D tables: which 0 Gauss points, 1 weights, 2 legendre at int [K][M],
   3 derivative [K][M], 4 nodal DOF locations [K], 5 grid DOF locations,
   6 PC-NSCBC matrix [M^2][M^2], 7 Lagrange derivative on the Gauss points
   [M][M] (row: point, column: basis), 8 the NSCBC coefficient
   -l'_0(r_0) of the nodal DOF basis (1 value). */
void orc_get_table(const orc_case *c, int which, double *out);

/* Test hooks: one sum-factorised tensor application (out = Oz Oy Ox in),
   one element mass-matrix solve, dU/dV at the points of an element. */
void orc_tensor_apply(int nix, int niy, int niz, int nox, int noy, int noz,
                      const double *ox, const double *oy, const double *oz,
                      const double *in, double *out);
int  orc_cg_solve(orc_case *c, int e, const double *rhs, double *x);
void orc_get_dudv(const orc_case *c, int e, double *out);

#ifdef __cplusplus
}
#endif
#endif
