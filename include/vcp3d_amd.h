/*
 * vcp3d_amd.h -- C ABI of the MI355X (gfx950) explicit DG residual and
 * SSP-RK3 synchronization step of VCP3D.
 *
 * The reference has no plugin layer: the seam is SolverClass
 * (VCP3D_CurviLinear/include/SolverClass.hpp:16-124). Each entry point below
 * names the reference member it replaces. All pointers are host pointers to
 * caller-owned memory, copied in or out; device memory, streams and kernels
 * stay inside the context. Every call returns 0 on success and a negative
 * code on failure (vcp_last_error() gives the text); nothing calls exit().
 *
 * Array layouts (FP64):
 *   elem_grid  [nElem][3][(n_poly_grid+1)^3]  the per-element grid DOFs
 *              (ElementClass::mCoorNodalGridDOFs, filled by
 *              SolverClass::ReadGrid, SolverClass.cpp:2988-3034)
 *   modal      [nElem][5][(P+1)^3]            entropy variables per element
 *              (ElementClass::mSol / mRes, unpadded)
 *   nodal      [nElem][(P+1)^3][5]            dimensional conservative
 *              variables at the nodal DOFs (SU2 restart payload,
 *              SolverClass.cpp:4375-4804)
 * Element index: i fastest, then j, then k.
 */
#ifndef VCP3D_AMD_H
#define VCP3D_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Input parameters of the path; names and meaning follow InputParamClass
   (VCP3D_CurviLinear/include/InputParamClass.hpp:18-301). */
typedef struct vcp_case_desc {
  int    n_elem[3];          /* elements in i, j, k (from the subface ranges)   */
  int    n_poly_grid;        /* "Polynomial degree grid"                        */
  int    n_poly_sol;         /* "Polynomial degree solution"                    */
  int    n_poly_int_rule;    /* "Polynomial degree integration rule"            */
  int    dof_location;       /* 1 equidistant, 2 LGL ("Location of the DOFs")   */
  int    periodic[3];        /* 1: Periodic1to1Trans pair on the i/j/k faces    */
  double mach;               /* "Mach number"                                   */
  double reynolds;           /* "Reynolds number"                               */
  double reynolds_length;    /* "Reynolds length"                               */
  double temp_free_stream;   /* "Free stream temperature"                       */
  double theta_sym;          /* "Theta parameter symmetrizing terms"            */
  double rel_penalty;        /* "Relative penalty parameter"                    */
  double cfl;                /* "CFL number"                                    */
  double dt_synchr;          /* "Synchronization time step" (dimensional)       */
  int    bug_compat;         /* 1: reproduce the reference's metric indexing in
                                the gradient conversion (metric of integration
                                point l used for variable l,
                                ElementClass.cpp:3865-3873, FluxesFace.cpp:
                                272-282, :325-332) and its slot order of the
                                transverse derivatives of a reconstructed
                                NSCBC state (BoundaryState.cpp:1261-1287);
                                0: use the point's metric and the matching
                                slots                                          */
  /* Physical block faces iMin, iMax, jMin, jMax, kMin, kMax (used where the
     direction is not periodic; one subface per block face), from the
     "Subface information" and "Prescribed boundary data" blocks
     (InputParamClass.cpp:1441-1576, :1340-1398). */
  int    bc[6];              /* VCP_BC_* */
  double bc_data[6][5];      /* dimensional prescribed data of the face:
                                NSCBC or subsonic outflow {Pressure}; static
                                NSCBC inflow {Density, VelocityX, VelocityY,
                                VelocityZ}; total NSCBC or subsonic inflow
                                {TotalPressure, TotalTemperature, VelocityDirX,
                                VelocityDirY, VelocityDirZ}; heat-flux wall
                                {HeatFlux}; isothermal wall {Temperature};
                                supersonic inflow {Density, VelocityX,
                                VelocityY, VelocityZ, Pressure}; inviscid wall
                                and supersonic outflow {}                        */
  /* "NSCBC tuning parameters" (InputParamClass.cpp:406-424) */
  double outlet_len;         /* "Outlet characteristic length"                  */
  double outlet_sigma;       /* "Outlet normal relaxation (sigma)"              */
  double outlet_beta_l;      /* "Outlet coupled transverse relaxation (beta_l)",
                                < 0: the boundary-averaged Mach number         */
  double outlet_beta_t;      /* "Outlet uncoupled transverse relaxation (beta_t)" */
  int    outlet_one_minus_beta;  /* "Outlet use one minus beta transverse relaxation" */
  int    outlet_average_local;   /* "Outlet use local (element) averaged Mach
                                    number": must be 0 (the reference keeps one
                                    value per subface object, written by every
                                    element, BoundaryState.cpp:985-990)         */
  double inlet_len;          /* "Inlet characteristic length"                   */
  double inlet_sigma;        /* "Inlet normal relaxation (sigma)"               */
  int    inlet_incoming;     /* "Inlet use outgoing wave amplitude"             */
  /* "Sponge layer", "Characteristic matching layer" and the "Sponge layer
     information" block (InputParamClass.cpp:428-432, :1196-1330) */
  int    sponge_layer;
  int    cml;
  int    sponge_elems[6];    /* layer thickness in elements per block face      */
  double damping_constant[6];
  double damping_exponent[6];
  /* "Monitor forces" and "Reference area force coefficients"
     (InputParamClass.cpp:79-80): force coefficients of the walls in
     monitoring[0..5], scaled by 2 / (gamma Mach^2 area) (SolverClass.cpp:
     3775-3783); reference_area <= 0 means the reference's default 1 */
  int    monitor_forces;
  double reference_area;
  /* "Compute average solution" (InputParamClass.cpp:365): running averages
     updated after every synchronization step (SolverClass.cpp:2415-2421) */
  int    compute_average;
  /* "Free stream velocity direction" (InputParamClass.cpp:214-216): the
     farfield state when the prescribed data do not fix the free stream
     (InputParamClass.cpp:991-1191) */
  double free_stream_dir[3];
  /* "Subgrid scale model" (InputParamClass.cpp:241-250): VCP_SGS_NONE,
     VCP_SGS_WALE or VCP_SGS_VREMAN (SGSModel.cpp:34-210); the eddy viscosity
     enters the volume and face viscous fluxes, the penalty and symmetrising
     terms, the time step, monitoring[7] and the averaged eddy viscosity of
     vcp_get_averages */
  int    sgs_model;
  /* "Body force per unit volume (x,y,z)" (InputParamClass.cpp:331-334),
     dimensional; body_force_set = 0 when the keyword is absent. The volume
     residual gains -w J (0, f, u.f) / p_ref at the integration points
     (ElementClass.cpp:4492-4533). */
  double body_force[3];
  int    body_force_set;
} vcp_case_desc;

enum { VCP_SGS_NONE = 0, VCP_SGS_WALE = 1, VCP_SGS_VREMAN = 2 };

/* Block-face boundary conditions ("Subface information",
   InputParamClass.cpp:1530-1548): BCOutflowNSCBC, BCInflowStaticNSCBC,
   BCInflowTotalNSCBC, BCInflowSubsonic (total conditions), BCOutflowSubsonic,
   BCWallHeatFlux, BCWallIsothermal, BCWallInviscid, BCInflowSupersonic,
   BCOutflowSupersonic, BCFarfield (the free stream). */
enum { VCP_BC_PERIODIC = 0, VCP_BC_OUTFLOW_NSCBC = 1, VCP_BC_INFLOW_STATIC_NSCBC = 2,
       VCP_BC_INFLOW_TOTAL_NSCBC = 3, VCP_BC_INFLOW_SUBSONIC = 4, VCP_BC_OUTFLOW_SUBSONIC = 5,
       VCP_BC_WALL_HEATFLUX = 6, VCP_BC_WALL_ISOTHERMAL = 7, VCP_BC_WALL_INVISCID = 8,
       VCP_BC_INFLOW_SUPERSONIC = 9, VCP_BC_OUTFLOW_SUPERSONIC = 10, VCP_BC_FARFIELD = 11 };

typedef struct vcp_ctx vcp_ctx;
typedef struct vcp_batch vcp_batch;

/* Setup: SolverClass::SolverClass + ReadGrid + ComputeMetricTerms
   (SolverClass.cpp:12-61, :2787-3034, :829-1016). */
int vcp_create(const vcp_case_desc *desc, const double *elem_grid, vcp_ctx **out);
int vcp_destroy(vcp_ctx *ctx);

/* Device memory the context holds (bytes of its allocations), no
   counterpart in the reference; the serving loop admits cases by it. */
int vcp_device_bytes(const vcp_ctx *ctx, size_t *bytes);

/* Sizes: number of elements, nDOFs1D (K), nInt1D (M). */
int vcp_sizes(const vcp_ctx *ctx, int *n_elem, int *n_dofs_1d, int *n_int_1d);
/* Reference conditions: uRef, pRef, rhoRef, non-dimensional mu
   (InputParamClass::ComputeReferenceConditions, InputParamClass.cpp:605-758). */
int vcp_reference(const vcp_ctx *ctx, double out[4]);

/* SolverClass::ReadRestartSolution (SolverClass.cpp:3040-3376): nodal
   dimensional conservative -> modal entropy variables. */
int vcp_set_restart(vcp_ctx *ctx, const double *nodal);
/* Sponge-layer target state: the averaged nodal primitive variables
   (rho, u, v, w, p), dimensional, [nElem][(P+1)^3][5]
   (SolverClass::ReadAveragedSolutionSpongeLayer, SpongeLayer.cpp:665-838,
   and ElementClass::InitializeDampingState, SpongeLayer.cpp:940-997).
   Needed before the first residual of a case with a sponge layer. */
int vcp_set_damping_state(vcp_ctx *ctx, const double *nodal_prim);
/* The same target from the case's own running averages, on the device (a
   case created with compute_average and a sponge layer, after at least one
   averaged synchronization step): the averaged primitive variables that
   SolverClass::ReadAveragedSolutionSpongeLayer reads from the averaged
   restart file (SpongeLayer.cpp:665-838) are taken from the device and
   converted as ElementClass::InitializeDampingState does (SpongeLayer.cpp:
   940-997), with no file or host copy in between. */
int vcp_damping_from_averages(vcp_ctx *ctx);
/* SolverClass::AverageDataNSCBC (SolverClass.cpp:2502-2581): boundary-averaged
   face-normal Mach number of the six block faces for the current state
   (0 where the face has no NSCBC). vcp_residual and vcp_sync_step evaluate it
   before every residual, as every reference call site of Residual does. */
int vcp_nscbc_average(vcp_ctx *ctx, double mach_avg[6]);
/* Direct access to the modal working variables (ElementClass::mSol). */
int vcp_set_solution(vcp_ctx *ctx, const double *modal);
int vcp_get_solution(vcp_ctx *ctx, double *modal);
/* Solution payload of SolverClass::WriteRestartSolution (SolverClass.cpp:4375-4560). */
int vcp_get_nodal(vcp_ctx *ctx, double *nodal);

/* SolverClass::Residual (SolverClass.cpp:3381-3795; SolverClass.hpp:72-73):
   residual multiplied by the inverse (entropy) mass matrix. monitoring[8]
   as the reference: 6 force coefficients, max Mach^2, max eddy viscosity. */
int vcp_residual(vcp_ctx *ctx, int compute_monitoring, double monitoring[8]);
/* No counterpart in the reference: vcp_residual for callers that cannot
   allocate device memory (code under HIP graph capture, serving loops,
   harnesses that forbid hipMalloc). It makes no hipMalloc/hipFree or other
   allocating call: the face-slot buffer that the face kernels fill and the
   volume kernel gathers (60 (P+1)^2 doubles per element) lives in a fixed
   64 MiB device array of the library, one per device, held until the
   process ends, cleared at the start of each call. Cases whose buffer
   exceeds it (e.g. more than 5592 elements at P = 4) are rejected with
   code -7. Every call on a device shares that array, so calls that overlap
   on one device (several host threads or streams) race. Results are
   identical to vcp_residual's. */
int vcp_residual_static(vcp_ctx *ctx, int compute_monitoring, double monitoring[8]);
/* ElementClass::mRes of every element after vcp_residual / vcp_sync_step. */
int vcp_get_residual(vcp_ctx *ctx, double *modal);

/* SolverClass::ExecuteTimeSynchronizationStep (SolverClass.cpp:2177-2422;
   SolverClass.hpp:50): SSP-RK3 steps until the synchronization time is
   reached, dt = min over elements, last step clipped by the 0.99 rule; the
   residual of the previous call is reused for the first stage. Returns the
   number of RK steps taken (>= 1) or a negative code. */
int vcp_sync_step(vcp_ctx *ctx, double monitoring[8]);

/* Execution option with no counterpart in the reference (the reference runs
   each stage as OpenMP loops, SolverClass.cpp:2200-2414): with enable != 0,
   vcp_sync_step replays every RK step (the stage-0 update, the three residual
   stages and, on the last step of a call, the next time step on a side
   stream) as an instantiated HIP graph, captured on first use. The stage
   kernels read the step's time step from device memory, so four captures
   (which state buffer is current x last step of the call or not) serve every
   step. Results are identical to the launch-by-launch path (enable = 0, the
   default). */
int vcp_set_graph(vcp_ctx *ctx, int enable);

/* No counterpart in the reference (measurement): durations of the face and
   volume launches inside vcp_sync_step, from HIP events recorded before,
   between and after them on the context's stream. enable = 1 starts
   recording and clears the sums, 0 stops it; the graph path (vcp_set_graph)
   is not timed, nor are stages past the 1024th of one vcp_sync_step call. Get: mean milliseconds per face launch (three directions)
   and per volume launch, and the number of timed stages. */
int vcp_set_kernel_timing(vcp_ctx *ctx, int enable);
int vcp_get_kernel_timing(vcp_ctx *ctx, double ms[2], int *stages);

/* The running averages (ElementClass::mAvePrim, mAveVelProd, mAveEddyVis;
   SolverClass::UpdateAverageData, SolverClass.cpp:3800-3850,
   ElementClass.cpp:3741-3804) of a case created with compute_average:
   avg [nElem][nDOF][12] dimensional (rho, u, v, w, p, uu, vv, ww, uv, uw, vw,
   eddy viscosity; 0 without an SGS model), the averaged columns of the reference's restart file;
   *n_steps the number of synchronization steps averaged. Set: to continue
   averaging from a restart ("Continue averaging from restart",
   SolverClass.cpp:3227-3240). */
int vcp_get_averages(vcp_ctx *ctx, double *avg, int *n_steps);
int vcp_set_averages(vcp_ctx *ctx, const double *avg, int n_steps);
/* The time-averaged force coefficients (SolverClass::mCfxAve, mCfyAve,
   mCfzAve, mCfxVisAve, mCfyVisAve, mCfzVisAve) of a case created with
   compute_average: updated from monitoring[0..5] after every averaged
   synchronization step (SolverClass.cpp:3858-3869), written as the first six
   restart meta data doubles (SolverClass.cpp:4407-4409) and read back from
   them when averaging continues from a restart (SolverClass.cpp:3236-3249). */
int vcp_get_force_averages(vcp_ctx *ctx, double cf[6]);
int vcp_set_force_averages(vcp_ctx *ctx, const double cf[6]);

/* ElementClass::ComputeTimeStep of every element (ElementClass.cpp:1522-1644). */
int vcp_time_steps(vcp_ctx *ctx, double *dt_elem);

/* No counterpart in the reference: the metric terms vcp_create computed on the
   device (ElementClass::ComputeMetricTerms, ElementClass.cpp:1027-1407;
   FinalMetric{I,J,K}Face, :1907-2088; ComputeLengthScales, :927-1022), for
   checks: vgeo [nElem][10][nInt] (J, dr/dx .. dt/dz at the integration
   points), fgeo [nElem][3][22][nInt2D] (per element and direction: unit
   normal and area factor of the min face seen from the neighbour below, that
   neighbour's inverse metric at its max face, the element's at its min face),
   econst [nElem][16] (length scales i, j, k, volume, averaged normals [3][3],
   LES length scale); any pointer may be NULL. */
int vcp_get_geometry(vcp_ctx *ctx, double *vgeo, double *fgeo, double *econst);
/* Iterations of the last per-element mass-matrix solve
   (CG_EntropyVarMassMatrix, CG_MassMatrix.cpp:250-504). */
int vcp_cg_iterations(vcp_ctx *ctx, int *iters);

/* Measurement: times `reps` residual evaluations with HIP events recorded on
   the context's stream around each kernel; writes the mean milliseconds per
   launch of [boundary + face kernels, volume+mass-matrix kernel] into ms[2]. */
int vcp_profile_residual(vcp_ctx *ctx, int reps, double ms[2]);

/* Execution option with no counterpart in the reference: the workgroup size
   of the volume kernel (threads per element; 0 restores the default). The
   sizes available for (nDOFs1D, nInt1D) come from vcp_volume_tiles, the
   default first; vcp_volume_tiles returns their number. Results of different
   sizes agree to rounding (the partial sums of the mass-matrix solve are
   grouped differently); a given size gives identical results every run. */
int vcp_set_tile(vcp_ctx *ctx, int nt);

/* Execution option with no counterpart in the reference (whose only precision
   switch is the all-FP32 su2double typedef, VCP3D_CurviLinear.hpp:63-118):
   VCP_PRECISION_MIXED runs the sum-factorisation contractions of the face and
   volume kernels and the mass-matrix solve in FP32 (the solve stops at
   ||r||_inf < 1e-5 ||r0||_inf), while the state and its element means, the
   pointwise flux physics, the face-slot gather, the metric terms, the NSCBC
   boundary kernels, the time step and the RK update stay in FP64.
   VCP_PRECISION_FP64 (the default) is the reference's arithmetic. */
enum { VCP_PRECISION_FP64 = 0, VCP_PRECISION_MIXED = 1 };
int vcp_set_precision(vcp_ctx *ctx, int mode);
int vcp_volume_tiles(int n_dofs_1d, int n_int_1d, int out[4]);

/* ---------------------------------------------------------------- batches
   Many independent cases stepped together, with no counterpart in the
   reference (one SolverClass per case, SolverClass.hpp:40-124): the cases
   are grouped by (nDOFs1D, nInt1D, bug_compat, precision), and each kernel of an RK
   stage is one launch per group over a work list of (case, element) items.
   Each case keeps its own time step and the synchronization-step clipping of
   ExecuteTimeSynchronizationStep (SolverClass.cpp:2356-2365) on the device:
   a case advanced n RK steps by the batch holds the state that n RK steps of
   vcp_sync_step give it alone with the same volume workgroup size. A case
   whose error flags are set (CG not converged, negative a^2, non-finite
   residual: where the reference calls Terminate) or whose step budget is
   spent is deactivated on the device; the others continue. */
int vcp_batch_create(vcp_batch **out);
/* Ends every membership (the contexts stay valid) and frees the batch. */
int vcp_batch_destroy(vcp_batch *b);
/* ctx joins with a budget of n_rk_steps RK steps (<= 0: unlimited). Its
   state must be set; its first residual is evaluated by the next
   vcp_batch_step. While it belongs to the batch, only the batch steps it:
   vcp_residual / vcp_sync_step / vcp_set_* on it fail, and the vcp_get_*
   calls wait for the batch's queued work. Cases that compute running
   averages cannot join. */
int vcp_batch_add(vcp_batch *b, vcp_ctx *ctx, int n_rk_steps);
/* ctx leaves the batch (after the batch's queued work), keeping its state;
   vcp_sync_step then continues from it with a new synchronization step. */
int vcp_batch_remove(vcp_batch *b, vcp_ctx *ctx);
/* Queues n_steps RK steps of every active case on the batch's stream and
   returns without waiting. */
int vcp_batch_step(vcp_batch *b, int n_steps);
int vcp_batch_wait(vcp_batch *b);
/* Waits, then writes the members (up to cap) and their status words
   status[i][5] = {active (the next RK step of the batch advances the case:
   no error flags and steps left), RK steps done, RK steps left, error flags,
   synchronization steps done}; returns the number of members. */
int vcp_batch_status(vcp_batch *b, vcp_ctx **ctxs, int *status, int cap);
/* Replay each RK step of the batch as one captured HIP graph (recaptured
   after a membership change). */
int vcp_batch_set_graph(vcp_batch *b, int enable);
/* Volume workgroup size of the group (nDOFs1D, nInt1D) (see vcp_set_tile). */
int vcp_batch_set_tile(vcp_batch *b, int n_dofs_1d, int n_int_1d, int nt);
/* Groups: info[g][7] = {nDOFs1D, nInt1D, bug_compat, cases, work items,
   volume workgroup size, precision}; returns the number of groups. */
int vcp_batch_groups(vcp_batch *b, int *info, int cap);
/* Measurement: mean milliseconds of `reps` residual evaluations per group,
   ms[2 g] boundary + face kernels, ms[2 g + 1] volume kernel (cap: entries
   of ms / 2); returns the number of groups. */
int vcp_batch_profile(vcp_batch *b, int reps, double *ms, int cap);

/* Host-only set-up (no device calls), for checking the tables and metrics
   without a GPU: SolverClass::ComputeMetricTerms (SolverClass.cpp:829-1016).
   vol [nElem][13][M^3] (J, dr/dx..dt/dz, x, y, z), faces [6][nElem][16][M^2]
   (iMin..kMax: unit normal, area factor, dr/dx..dt/dz, x, y, z), lens
   [nElem][4] (length scales I, J, K, volume), nbr [nElem][6] (-1 boundary).
   Any output pointer may be NULL. */
int vcp_host_setup(const vcp_case_desc *desc, const double *elem_grid, double *vol,
                   double *faces, double *lens, int *nbr);

const char *vcp_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
