// C ABI of the GPU path (include/vcp3d_amd.h): context set-up, device
// buffers and the host sequencing of SolverClass::Residual and
// SolverClass::ExecuteTimeSynchronizationStep.
#include "ctx.hpp"
#include "geometry.hpp"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <utility>
#include <vector>

namespace vcp {

namespace {
thread_local std::string g_error;
}

int fail(const std::string &msg, int code)
{
  g_error = msg;
  return code;
}

}  // namespace vcp

using vcp::fail;

namespace {

// device scalars read back with the monitoring data, and the slot of the
// error flags in the pinned host buffer
constexpr int kScalRead = vcp::kScalEddy + 1, kHostFlags = vcp::kScalCount;

size_t state_size(const vcp_ctx *c) { return (size_t) c->dc.nElem * 5 * c->dc.nDOF; }

int flag_error(int flags)
{
  if (flags & vcp::kFlagNegativeA2) return fail("Negative square speed of sound", -4);
  if (flags & vcp::kFlagCGNotConverged) return fail("Convergence not reached for CG algorithm", -5);
  if (flags & vcp::kFlagNonFinite) return fail("non-finite residual", -6);
  return 0;
}

int check_flags(vcp_ctx *c)
{
  int flags = 0;
  HIP_TRY(hipMemcpyAsync(&flags, c->dc.flags, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return flag_error(flags);
}


int residual_pass(vcp_ctx *c, const vcp::StageCtl &ctl)
{
  // timed: the same launches with an event before, between and after the
  // face and volume launches, for at most 1024 stages per call (the events
  // are read back at the end of the call)
  if (!c->kt.on || c->use_graph || c->kt.used >= 3 * 1024) return vcp::residual_pass(c, ctl, c->stream);
  if (c->prm.sponge_layer && !c->damping_set)
    return fail("sponge layer damping state not set (vcp_set_damping_state)");
  vcp_ctx::KernelTiming &kt = c->kt;
  while (kt.ev.size() < kt.used + 3) {
    hipEvent_t e = nullptr;
    HIP_TRY(hipEventCreate(&e));
    kt.ev.push_back(e);
  }
  hipEvent_t *ev = kt.ev.data() + kt.used;
  HIP_TRY(vcp::launch_boundary(c->dc, c->stream, ctl.monitor != 0));
  HIP_TRY(hipEventRecord(ev[0], c->stream));
  HIP_TRY(vcp::launch_faces(c->dc, c->stream));
  HIP_TRY(hipEventRecord(ev[1], c->stream));
  HIP_TRY(vcp::launch_volume(c->dc, ctl, c->stream));
  HIP_TRY(hipEventRecord(ev[2], c->stream));
  kt.used += 3;   // counted once all three are recorded
  return 0;
}

// adds the recorded launches' durations to the sums (after a synchronisation
// of the stream they were recorded on)
int collect_timing(vcp_ctx *c)
{
  vcp_ctx::KernelTiming &kt = c->kt;
  for (size_t i = 0; i + 3 <= kt.used; i += 3) {
    float f = 0.f, v = 0.f;
    HIP_TRY(hipEventElapsedTime(&f, kt.ev[i], kt.ev[i + 1]));
    HIP_TRY(hipEventElapsedTime(&v, kt.ev[i + 1], kt.ev[i + 2]));
    kt.face_ms += f; kt.vol_ms += v; ++kt.launches;
  }
  kt.used = 0;
  return 0;
}

// a context that a batch steps is driven by the batch only
int not_in_batch(const vcp_ctx *c)
{
  return c->batch ? fail("the context belongs to a batch (vcp_batch_remove it first)") : 0;
}

// The stages of one SSP-RK3 step (Shu-Osher form, SolverClass.cpp:2192-2194,
// :2385-2393), with U0 holding the step's state; its time step dt is passed to
// the stage-0 kernel, which stores it to scal[kScalDt] for the other stages,
// or (dt null, for a captured step) is already in scal[kScalDt]. Stage 0 uses the residual of the previous evaluation (first same as
// last); U is final after the second stage, so on the last step of a call the
// next call's time step runs on the side stream beside the third stage, which
// only evaluates the residual, and is joined at the end.
int rk_stages(vcp_ctx *c, bool last, const double *dt)
{
  static const double alphaRK[] = {1.0, 3.0 / 4.0, 1.0 / 3.0};
  static const double betaRK[] = {1.0, 1.0 / 4.0, 2.0 / 3.0};
  int rc = 0;
  HIP_TRY(vcp::launch_stage0(c->dc, dt, c->stream));
  const vcp::StageCtl s1{0, 1, alphaRK[1], betaRK[1], 0};
  const vcp::StageCtl s2{0, 1, alphaRK[2], betaRK[2], 0};
  const vcp::StageCtl s3{1, 0, 0.0, 0.0, last ? 1 : 0};
  if ((rc = residual_pass(c, s1))) return rc;
  if ((rc = residual_pass(c, s2))) return rc;
  if (last) {
    HIP_TRY(hipEventRecord(c->evState, c->stream));
    HIP_TRY(hipStreamWaitEvent(c->side, c->evState, 0));
    HIP_TRY(vcp::launch_time_step(c->dc, c->side));
    HIP_TRY(hipEventRecord(c->evDt, c->side));
  }
  if ((rc = residual_pass(c, s3))) return rc;
  if (last) HIP_TRY(hipStreamWaitEvent(c->stream, c->evDt, 0));
  return 0;
}

// rk_stages replayed from an instantiated graph. The launches of a step
// depend only on which state buffer is current and on whether it is the last
// step of the call (the time step is read from device memory, where a kernel
// launched before the graph puts it), so four captures cover every step.
int rk_graph(vcp_ctx *c, bool last, double dt)
{
  HIP_TRY(vcp::launch_set_dt(c->dc, dt, c->stream));
  hipGraphExec_t &exec = c->graph[c->dc.U == c->U_first ? 0 : 1][last ? 1 : 0];
  if (!exec) {
    hipGraph_t g = nullptr;
    HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    const int rc = rk_stages(c, last, nullptr);
    const hipError_t he = hipStreamEndCapture(c->stream, &g);
    if (rc || he != hipSuccess) {
      if (g) (void) hipGraphDestroy(g);
      return rc ? rc : fail(std::string("hipStreamEndCapture: ") + hipGetErrorString(he), -3);
    }
    const hipError_t hi = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
    (void) hipGraphDestroy(g);
    if (hi != hipSuccess) {
      exec = nullptr;
      return fail(std::string("hipGraphInstantiate: ") + hipGetErrorString(hi), -3);
    }
  }
  HIP_TRY(hipGraphLaunch(exec, c->stream));
  return 0;
}

// vcp_case_desc -> CaseParams (field by field, same meaning).
void params_from_desc(const vcp_case_desc *d, vcp::CaseParams &p)
{
  for (int i = 0; i < 3; ++i) { p.n_elem[i] = d->n_elem[i]; p.periodic[i] = d->periodic[i]; }
  p.n_poly_grid = d->n_poly_grid; p.n_poly_sol = d->n_poly_sol; p.n_poly_int_rule = d->n_poly_int_rule;
  p.dof_location = d->dof_location; p.mach = d->mach; p.reynolds = d->reynolds;
  p.reynolds_length = d->reynolds_length; p.temp_free_stream = d->temp_free_stream;
  p.theta_sym = d->theta_sym; p.rel_penalty = d->rel_penalty; p.cfl = d->cfl; p.dt_synchr = d->dt_synchr;
  p.bug_compat = d->bug_compat;
  for (int f = 0; f < 6; ++f) {
    p.bc[f] = d->bc[f];
    for (int q = 0; q < 5; ++q) p.bc_data[f][q] = d->bc_data[f][q];
    p.sponge_elems[f] = d->sponge_elems[f];
    p.damping_constant[f] = d->damping_constant[f];
    p.damping_exponent[f] = d->damping_exponent[f];
  }
  p.monitor_forces = d->monitor_forces;
  p.compute_average = d->compute_average;
  for (int i = 0; i < 3; ++i) p.free_stream_dir[i] = d->free_stream_dir[i];
  p.reference_area = d->reference_area > 0.0 ? d->reference_area : 1.0;
  p.outlet_len = d->outlet_len; p.outlet_sigma = d->outlet_sigma;
  p.outlet_beta_l = d->outlet_beta_l; p.outlet_beta_t = d->outlet_beta_t;
  p.outlet_one_minus_beta = d->outlet_one_minus_beta; p.outlet_average_local = d->outlet_average_local;
  p.inlet_len = d->inlet_len; p.inlet_sigma = d->inlet_sigma; p.inlet_incoming = d->inlet_incoming;
  p.sponge_layer = d->sponge_layer; p.cml = d->cml;
  p.sgs_model = d->sgs_model;
  p.body_force_set = d->body_force_set;
  for (int i = 0; i < 3; ++i) p.body_force[i] = d->body_force_set ? d->body_force[i] : 0.0;
}

// Monitoring data of SolverClass::Residual (SolverClass.cpp:3754-3783) from
// the device scalars sc[0..kScalForce+6): force coefficients = wall force sums
// times 2 / (gamma Mach^2 area), maximum Mach^2, maximum eddy viscosity (0: no
// SGS model).
void fill_monitoring(const vcp_ctx *c, double md[8], bool mon, const double *sc)
{
  const double area = c->prm.reference_area > 0.0 ? c->prm.reference_area : 1.0;
  const double fc = 2.0 / (1.4 * c->prm.mach * c->prm.mach * area);
  for (int i = 0; i < 6; ++i) md[i] = mon && c->dc.monitor_forces ? sc[vcp::kScalForce + i] * fc : 0.0;
  md[6] = mon ? sc[1] : 0.0;
  md[7] = mon ? (c->dc.sgs ? sc[vcp::kScalEddy] : 0.0) : -1.0;
}

// Metric terms of every element on the device (launch_geometry) into the
// context's vgeo, fgeo and econst: the grid and the tables are uploaded to
// temporary buffers, freed before returning; the LES length scale
// volume^(1/3) / P of an SGS case is set on the host from the volumes.
int device_geometry(vcp_ctx *c, const vcp::CaseParams &p, const double *elem_grid)
{
  vcp::DeviceCase &dc = c->dc;
  const int Kg = p.n_poly_grid + 1, nG = Kg * Kg * Kg;
  const std::vector<double> tab = vcp::geometry_tables(c->tab);
  double *dGrid = nullptr, *dTab = nullptr;
  int *dFlag = nullptr, flag = 0;
  const size_t nGrid = (size_t) dc.nElem * 3 * nG;
  hipError_t he = hipMalloc((void **) &dGrid, nGrid * sizeof(double));
  if (he == hipSuccess) he = hipMalloc((void **) &dTab, tab.size() * sizeof(double));
  if (he == hipSuccess) he = hipMalloc((void **) &dFlag, sizeof(int));
  if (he == hipSuccess) he = hipMemcpy(dGrid, elem_grid, nGrid * sizeof(double), hipMemcpyHostToDevice);
  if (he == hipSuccess) he = hipMemcpy(dTab, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice);
  if (he == hipSuccess) he = hipMemset(dFlag, 0, sizeof(int));
  if (he == hipSuccess)
    he = hipMemset(dc.fgeo, 0, (size_t) dc.nElem * 3 * vcp::kFaceGeo * dc.nInt2D * sizeof(double));
  if (he == hipSuccess) {
    vcp::GeoArgs a{dGrid, dTab, dc.nbr, dc.vgeo, dc.fgeo, dc.econst, dFlag, dc.nElem, Kg};
    he = vcp::launch_geometry(a, dc.M, c->stream);
  }
  if (he == hipSuccess) he = hipStreamSynchronize(c->stream);
  if (he == hipSuccess) he = hipMemcpy(&flag, dFlag, sizeof(int), hipMemcpyDeviceToHost);
  if (dGrid) (void) hipFree(dGrid);
  if (dTab) (void) hipFree(dTab);
  if (dFlag) (void) hipFree(dFlag);
  if (he != hipSuccess) return fail(std::string("device metric terms: ") + hipGetErrorString(he), -3);
  if (flag) return fail("Negative or zero Jacobians encountered");
  if (dc.sgs) {
    std::vector<double> ec((size_t) dc.nElem * vcp::kElemConst);
    HIP_TRY(hipMemcpy(ec.data(), dc.econst, ec.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int e = 0; e < dc.nElem; ++e)
      ec[(size_t) e * vcp::kElemConst + vcp::kElemLenLES] = std::pow(ec[(size_t) e * vcp::kElemConst + 3], 1.0 / 3.0) / p.n_poly_sol;
    HIP_TRY(hipMemcpy(dc.econst, ec.data(), ec.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  return 0;
}

}  // namespace

namespace vcp {

int residual_pass(vcp_ctx *c, const StageCtl &ctl, hipStream_t s)
{
  if (c->prm.sponge_layer && !c->damping_set)
    return fail("sponge layer damping state not set (vcp_set_damping_state)");
  HIP_TRY(launch_boundary(c->dc, s, ctl.monitor != 0));
  HIP_TRY(launch_faces(c->dc, s));
  HIP_TRY(launch_volume(c->dc, ctl, s));
  return 0;
}

}  // namespace vcp

extern "C" {

const char *vcp_last_error(void) { return vcp::g_error.c_str(); }

int vcp_create(const vcp_case_desc *d, const double *elem_grid, vcp_ctx **out)
{
  VCP_GUARD_BEGIN
  if (!d || !elem_grid || !out) return fail("null argument");
  *out = nullptr;
  vcp::CaseParams p;
  params_from_desc(d, p);
  if (p.n_elem[0] <= 0 || p.n_elem[1] <= 0 || p.n_elem[2] <= 0) return fail("invalid element counts");
  if (p.n_poly_sol <= 0 || p.n_poly_sol > 9 || p.n_poly_grid <= 0 || p.n_poly_grid > 9)
    return fail("polynomial degrees must be in 1..9");
  if (p.n_poly_int_rule < 2 * p.n_poly_sol)
    return fail("\"Polynomial degree integration rule\" not accurate enough");
  {
    const std::string why = vcp::check_case(p);
    if (!why.empty()) return fail(why);
  }
  const int K = p.n_poly_sol + 1, M = vcp::n_int1d(p.n_poly_int_rule);
  if (!vcp::supported_pair(K, M)) return fail("unsupported (nDOFs1D, nInt1D) pair");

  std::unique_ptr<vcp_ctx> holder(new vcp_ctx());
  vcp_ctx *c = holder.get();
  c->prm = p;
  vcp::build_tables(p, c->tab);
  // host: the spacing check, neighbours and the metric terms of the elements
  // the boundary and sponge set-ups read; the device computes all elements'
  // metric terms below
  vcp::Geometry g;
  const std::string err = vcp::build_geometry(p, c->tab, elem_grid, g, false);
  if (!err.empty()) return fail(err);
  c->rc = vcp::reference_conditions(p);
  c->dt_sync = p.dt_synchr * c->rc.u_ref;
  vcp::OpenBoundary ob;
  std::string serr = vcp::setup_open_boundary(p, c->tab, g, c->rc, ob);
  if (serr.empty()) serr = vcp::setup_sponge(p, c->tab, g, elem_grid, c->sponge);
  if (!serr.empty()) return fail(serr);

  vcp::DeviceCase &dc = c->dc;
  const int nElem = p.n_elem[0] * p.n_elem[1] * p.n_elem[2];
  dc.nElem = nElem; dc.K = K; dc.M = M; dc.P = p.n_poly_sol;
  dc.nDOF = K * K * K; dc.nInt = M * M * M; dc.nInt2D = M * M;
  dc.bug_compat = p.bug_compat; dc.mu = c->rc.mu; dc.theta_sym = p.theta_sym; dc.rel_penalty = p.rel_penalty;
  dc.cfl = p.cfl; dc.u_ref = c->rc.u_ref; dc.p_ref = c->rc.p_ref; dc.rho_ref = c->rc.rho_ref;
  dc.sgs = p.sgs_model; dc.nAvg = dc.sgs ? 12 : 11;
  // non-dimensional body force (ElementClass.cpp:4498-4502)
  dc.fbodyOn = p.body_force_set != 0;
  const double pRefInv = 1.0 / dc.p_ref;
  for (int i = 0; i < 3; ++i) dc.fbody[i] = pRefInv * p.body_force[i];

  // table block
  const vcp::Tables &t = c->tab;
  std::vector<double> tb;
  auto put = [&](const std::vector<double> &v) { const int off = (int) tb.size(); tb.insert(tb.end(), v.begin(), v.end()); return off; };
  dc.to.A = put(t.A); dc.to.D = put(t.D); dc.to.AT = put(t.AT); dc.to.DT = put(t.DT); dc.to.LD = put(t.LD);
  dc.to.DLD = put(t.DLD);
  dc.to.fmin = put(t.f_min); dc.to.fmax = put(t.f_max); dc.to.dfmin = put(t.df_min); dc.to.dfmax = put(t.df_max);
  for (int i = 0; i < M * K; ++i) { dc.opA[i] = t.A[i]; dc.opD[i] = t.D[i]; }
  for (int q = 0; q < K; ++q) {
    dc.opF[0][q] = t.f_min[q]; dc.opF[1][q] = t.f_max[q]; dc.opF[2][q] = t.df_min[q]; dc.opF[3][q] = t.df_max[q];
  }
  for (int q = 0; q < M; ++q) dc.opW[q] = t.w_int[q];
  for (int q = 0; q < K; ++q) dc.opC0[q] = t.c0[q];
  dc.to.w3 = put(t.w3); dc.to.w2 = put(t.w2); dc.to.bc = put(t.basis_center);
  dc.to.P1 = put(t.P1); dc.to.dLq = put(t.dLq); dc.to.Lq = put(t.Lq); dc.to.total = (int) tb.size();
  // open boundaries: face list, own face metric, parameters
  dc.nBFace = (int) ob.faces.size() / 2;
  std::vector<double> bgeo((size_t) dc.nBFace * 13 * M * M);
  for (int b = 0; b < dc.nBFace; ++b) {
    const int e = ob.faces[2 * b], f = ob.faces[2 * b + 1];
    std::memcpy(&bgeo[(size_t) b * 13 * M * M], &g.face[f][(size_t) e * 16 * M * M], sizeof(double) * 13 * M * M);
  }
  vcp::NscbcParams &nb = dc.nscbc;
  nb.outlet_len = p.outlet_len; nb.outlet_sigma = p.outlet_sigma;
  nb.outlet_beta_l = p.outlet_beta_l; nb.outlet_beta_t = p.outlet_beta_t;
  nb.inlet_len = p.inlet_len; nb.inlet_sigma = p.inlet_sigma; nb.dell0 = t.dell0;
  nb.outlet_one_minus_beta = p.outlet_one_minus_beta; nb.inlet_incoming = p.inlet_incoming;
  for (int f = 0; f < 6; ++f) {
    nb.bc[f] = p.periodic[f / 2] ? vcp::kBCPeriodic : p.bc[f];
    for (int q = 0; q < 5; ++q) nb.data[f][q] = ob.data[f][q];
    nb.invArea[f] = ob.area[f] > 0.0 ? 1.0 / ob.area[f] : 0.0;
  }
  for (int f = 0; f < 7; ++f) nb.begin[f] = ob.begin[f];
  // "Monitor forces" is switched off without walls (InputParamClass.cpp:1572-1573)
  dc.monitor_forces = 0;
  for (int f = 0; f < 6; ++f)
    if (p.monitor_forces && !p.periodic[f / 2] && vcp::bc_is_wall(p.bc[f])) dc.monitor_forces = 1;
  dc.nSponge = c->sponge.n; dc.cml = p.cml;

  const int nInt = dc.nInt, n2 = dc.nInt2D;
  const size_t nVgeo = (size_t) nElem * vcp::kVolGeo * nInt, nFgeo = (size_t) nElem * 3 * vcp::kFaceGeo * n2;
  const size_t nEcon = (size_t) nElem * vcp::kElemConst;
  hipError_t he = hipStreamCreate(&c->stream);
  if (he == hipSuccess) he = hipHostMalloc((void **) &c->host, (kHostFlags + 1) * sizeof(double), hipHostMallocDefault);
  if (he == hipSuccess) he = hipStreamCreate(&c->side);
  if (he == hipSuccess) he = hipEventCreateWithFlags(&c->evState, hipEventDisableTiming);
  if (he == hipSuccess) he = hipEventCreateWithFlags(&c->evDt, hipEventDisableTiming);
  const size_t ns = (size_t) nElem * 5 * dc.nDOF;
  if (he == hipSuccess) he = c->alloc(&dc.tables, tb.size());
  if (he == hipSuccess) he = c->alloc(&dc.vgeo, nVgeo);
  if (he == hipSuccess) he = c->alloc(&dc.fgeo, nFgeo);
  if (he == hipSuccess) he = c->alloc(&dc.econst, nEcon);
  if (he == hipSuccess) he = c->alloc(&dc.nbr, g.nbr.size());
  if (he == hipSuccess) he = c->alloc(&dc.U, ns);
  if (he == hipSuccess) he = c->alloc(&dc.U0, ns);
  c->U_first = dc.U;
  if (he == hipSuccess) he = c->alloc(&dc.R, ns);
  // face slots: per element and block face, the value and normal-derivative
  // coefficient planes [2][5][K][K] of the lifted face residual
  const size_t nslot = (size_t) nElem * 6 * 10 * dc.K * dc.K;
  if (he == hipSuccess) he = c->alloc(&dc.faceRes, nslot);
  if (he == hipSuccess) he = c->alloc(&dc.scal, vcp::kScalCount);
  if (he == hipSuccess) he = c->alloc(&dc.flags, 4);
  if (he == hipSuccess) he = c->alloc(&dc.cgIter, (size_t) nElem);
  if (he == hipSuccess) he = c->alloc(&dc.dtElem, (size_t) nElem);
  if (he == hipSuccess) he = c->alloc(&dc.mach2Elem, (size_t) nElem);
  if (he == hipSuccess && p.compute_average) {
    he = c->alloc(&dc.avg, (size_t) nElem * dc.nAvg * dc.nDOF);
    if (he == hipSuccess) he = hipMemset(dc.avg, 0, (size_t) nElem * dc.nAvg * dc.nDOF * sizeof(double));
  }
  if (he == hipSuccess && dc.sgs) {
    // nodal metric and the eddy-viscosity maxima (entries of faces that are
    // never written stay 0)
    he = c->alloc(&dc.vgeoN, g.nodal.size());
    if (he == hipSuccess)
      he = hipMemcpy(dc.vgeoN, g.nodal.data(), g.nodal.size() * sizeof(double), hipMemcpyHostToDevice);
    const size_t ne = (size_t) 4 * nElem + (size_t) ob.faces.size() / 2;
    if (he == hipSuccess) he = c->alloc(&dc.eddyPart, ne);
    if (he == hipSuccess) he = hipMemset(dc.eddyPart, 0, ne * sizeof(double));
  }
  if (he == hipSuccess && dc.nBFace) {
    he = c->alloc(&dc.bface, ob.faces.size());
    if (he == hipSuccess) he = c->alloc(&dc.bgeo, bgeo.size());
    if (he == hipSuccess) he = c->alloc(&dc.machPart, (size_t) dc.nBFace);
    if (he == hipSuccess) he = c->alloc(&dc.forcePart, (size_t) dc.nBFace * 6);
    if (he == hipSuccess) he = hipMemcpy(dc.bface, ob.faces.data(), ob.faces.size() * sizeof(int), hipMemcpyHostToDevice);
    if (he == hipSuccess) he = hipMemcpy(dc.bgeo, bgeo.data(), bgeo.size() * sizeof(double), hipMemcpyHostToDevice);
  }
  if (he == hipSuccess && dc.nSponge) {
    const vcp::Sponge &sp = c->sponge;
    const size_t ns5 = (size_t) dc.nSponge * 5 * nInt;
    he = c->alloc(&dc.spIdx, sp.idx.size());
    if (he == hipSuccess) he = c->alloc(&dc.spCoef, sp.coef.size());
    if (he == hipSuccess) he = c->alloc(&dc.dampFn, sp.fn.size());
    if (he == hipSuccess) he = c->alloc(&dc.dampState, ns5);
    if (he == hipSuccess) he = c->alloc(&dc.srcBuf, ns5);
    if (he == hipSuccess) he = hipMemcpy(dc.spIdx, sp.idx.data(), sp.idx.size() * sizeof(int), hipMemcpyHostToDevice);
    if (he == hipSuccess) he = hipMemcpy(dc.spCoef, sp.coef.data(), sp.coef.size() * sizeof(double), hipMemcpyHostToDevice);
    if (he == hipSuccess) he = hipMemcpy(dc.dampFn, sp.fn.data(), sp.fn.size() * sizeof(double), hipMemcpyHostToDevice);
    if (he == hipSuccess) he = hipMemset(dc.dampState, 0, ns5 * sizeof(double));
  }
  if (he == hipSuccess && dc.fbodyOn) he = c->alloc(&dc.bfBuf, (size_t) nElem * 4 * nInt);
  if (he == hipSuccess) he = hipMemcpy(dc.tables, tb.data(), tb.size() * sizeof(double), hipMemcpyHostToDevice);
  if (he == hipSuccess) he = hipMemcpy(dc.nbr, g.nbr.data(), g.nbr.size() * sizeof(int), hipMemcpyHostToDevice);
  if (he != hipSuccess) return fail(std::string("device allocation/upload: ") + hipGetErrorString(he), -3);
  if (int rc_ = device_geometry(c, p, elem_grid)) return rc_;
  if (he == hipSuccess) he = hipMemset(dc.U, 0, ns * sizeof(double));
  if (he == hipSuccess) he = hipMemset(dc.R, 0, ns * sizeof(double));
  if (he == hipSuccess) he = hipMemset(dc.faceRes, 0, nslot * sizeof(double));
  if (he == hipSuccess) he = hipMemset(dc.flags, 0, 4 * sizeof(int));
  if (he == hipSuccess) he = hipMemset(dc.scal, 0, vcp::kScalCount * sizeof(double));
  // a case is active outside a batch (the batch kernels skip inactive cases)
  if (he == hipSuccess) {
    const int active = 1;
    he = hipMemcpy(dc.flags + vcp::kWordActive, &active, sizeof(int), hipMemcpyHostToDevice);
  }
  if (he != hipSuccess) {
    return fail(std::string("device set-up failed: ") + hipGetErrorString(he), -3);
  }
  *out = holder.release();
  return 0;
  VCP_GUARD_END
}

int vcp_host_setup(const vcp_case_desc *d, const double *elem_grid, double *vol, double *faces,
                   double *lens, int *nbr)
{
  VCP_GUARD_BEGIN
  if (!d || !elem_grid) return fail("null argument");
  vcp::CaseParams p;
  params_from_desc(d, p);
  if (p.n_elem[0] <= 0 || p.n_elem[1] <= 0 || p.n_elem[2] <= 0) return fail("invalid element counts");
  if (p.n_poly_sol <= 0 || p.n_poly_sol > 9 || p.n_poly_grid <= 0 || p.n_poly_grid > 9)
    return fail("polynomial degrees must be in 1..9");
  const std::string chk = vcp::check_case(p);   // the same case checks as vcp_create
  if (!chk.empty()) return fail(chk);
  vcp::Tables t;
  vcp::build_tables(p, t);
  vcp::Geometry g;
  const std::string err = vcp::build_geometry(p, t, elem_grid, g);
  if (!err.empty()) return fail(err);
  if (vol) std::memcpy(vol, g.vol.data(), g.vol.size() * sizeof(double));
  if (faces)
    for (int f = 0; f < 6; ++f)
      std::memcpy(faces + f * g.face[f].size(), g.face[f].data(), g.face[f].size() * sizeof(double));
  if (lens) std::memcpy(lens, g.lens.data(), g.lens.size() * sizeof(double));
  if (nbr) std::memcpy(nbr, g.nbr.data(), g.nbr.size() * sizeof(int));
  return 0;
  VCP_GUARD_END
}

int vcp_destroy(vcp_ctx *c)
{
  VCP_GUARD_BEGIN
  if (!c) return 0;
  int rc = 0;
  if (c->batch) rc = vcp_batch_remove(c->batch, c);
  (void) hipStreamSynchronize(c->stream);
  delete c;
  return rc;
  VCP_GUARD_END
}

int vcp_sizes(const vcp_ctx *c, int *n_elem, int *k, int *m)
{
  VCP_GUARD_BEGIN
  if (!c) return fail("null context");
  if (n_elem) *n_elem = c->dc.nElem;
  if (k) *k = c->dc.K;
  if (m) *m = c->dc.M;
  return 0;
  VCP_GUARD_END
}

int vcp_reference(const vcp_ctx *c, double out[4])
{
  VCP_GUARD_BEGIN
  if (!c || !out) return fail("null argument");
  out[0] = c->rc.u_ref; out[1] = c->rc.p_ref; out[2] = c->rc.rho_ref; out[3] = c->rc.mu;
  return 0;
  VCP_GUARD_END
}

int vcp_set_restart(vcp_ctx *c, const double *nodal)
{
  VCP_GUARD_BEGIN
  if (!c || !nodal) return fail("null argument");
  if (int rc_ = not_in_batch(c)) return rc_;
  const int nD = c->dc.nDOF;
  std::vector<double> modal(state_size(c));
#pragma omp parallel for schedule(static)
  for (int e = 0; e < c->dc.nElem; ++e)
    vcp::nodal_cons_to_modal_entropy(c->tab, c->rc, nodal + (size_t) e * nD * 5, &modal[(size_t) e * 5 * nD]);
  return vcp_set_solution(c, modal.data());
  VCP_GUARD_END
}

int vcp_set_damping_state(vcp_ctx *c, const double *nodal_prim)
{
  VCP_GUARD_BEGIN
  if (!c || !nodal_prim) return fail("null argument");
  if (int rc_ = not_in_batch(c)) return rc_;
  if (!c->prm.sponge_layer) return fail("no sponge layer in this case");
  const vcp::Sponge &sp = c->sponge;
  const int nD = c->dc.nDOF, nInt = c->dc.nInt;
  std::vector<double> ds((size_t) sp.n * 5 * nInt);
#pragma omp parallel for schedule(static)
  for (int e = 0; e < c->dc.nElem; ++e)
    if (sp.idx[e] >= 0)
      vcp::damping_state(c->tab, c->rc, nodal_prim + (size_t) e * nD * 5, &ds[(size_t) sp.idx[e] * 5 * nInt]);
  HIP_TRY(hipMemcpyAsync(c->dc.dampState, ds.data(), ds.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->damping_set = true;
  c->res_valid = false;
  return 0;
  VCP_GUARD_END
}

int vcp_damping_from_averages(vcp_ctx *c)
{
  VCP_GUARD_BEGIN
  if (!c) return fail("null context");
  if (int rc_ = not_in_batch(c)) return rc_;
  if (!c->prm.sponge_layer) return fail("no sponge layer in this case");
  if (!c->dc.avg) return fail("the case was created without compute_average");
  if (c->n_avg < 1) return fail("no averaged synchronization step yet");
  HIP_TRY(vcp::launch_damping_from_averages(c->dc, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->damping_set = true;
  c->res_valid = false;
  return 0;
  VCP_GUARD_END
}

int vcp_nscbc_average(vcp_ctx *c, double mach_avg[6])
{
  VCP_GUARD_BEGIN
  if (!c || !mach_avg) return fail("null argument");
  if (int rc_ = not_in_batch(c)) return rc_;
  for (int f = 0; f < 6; ++f) mach_avg[f] = 0.0;
  if (!c->dc.nBFace) return 0;
  HIP_TRY(vcp::launch_boundary(c->dc, c->stream));
  std::vector<double> part(c->dc.nBFace);
  HIP_TRY(hipMemcpyAsync(part.data(), c->dc.machPart, part.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  for (int f = 0; f < 6; ++f) {
    double sum = 0.0;
    for (int b = c->dc.nscbc.begin[f]; b < c->dc.nscbc.begin[f + 1]; ++b) sum += part[b];
    mach_avg[f] = sum * c->dc.nscbc.invArea[f];
  }
  return 0;
  VCP_GUARD_END
}

int vcp_set_solution(vcp_ctx *c, const double *modal)
{
  VCP_GUARD_BEGIN
  if (!c || !modal) return fail("null argument");
  if (int rc_ = not_in_batch(c)) return rc_;
  HIP_TRY(hipMemcpyAsync(c->dc.U, modal, state_size(c) * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->res_valid = false;
  c->dt_ready = false;
  return 0;
  VCP_GUARD_END
}

int vcp_get_solution(vcp_ctx *c, double *modal)
{
  VCP_GUARD_BEGIN
  if (!c || !modal) return fail("null argument");
  if (int rc_ = vcp::batch_quiesce(c)) return rc_;
  HIP_TRY(hipMemcpyAsync(modal, c->dc.U, state_size(c) * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return 0;
  VCP_GUARD_END
}

int vcp_get_nodal(vcp_ctx *c, double *nodal)
{
  VCP_GUARD_BEGIN
  if (!c || !nodal) return fail("null argument");
  if (int rc_ = vcp::batch_quiesce(c)) return rc_;
  std::vector<double> modal(state_size(c));
  const int rc = vcp_get_solution(c, modal.data());
  if (rc) return rc;
  const int nD = c->dc.nDOF;
#pragma omp parallel for schedule(static)
  for (int e = 0; e < c->dc.nElem; ++e)
    vcp::modal_entropy_to_nodal_cons(c->tab, c->rc, &modal[(size_t) e * 5 * nD], nodal + (size_t) e * nD * 5);
  return 0;
  VCP_GUARD_END
}

int vcp_get_residual(vcp_ctx *c, double *modal)
{
  VCP_GUARD_BEGIN
  if (!c || !modal) return fail("null argument");
  if (int rc_ = vcp::batch_quiesce(c)) return rc_;
  HIP_TRY(hipMemcpyAsync(modal, c->dc.R, state_size(c) * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return 0;
  VCP_GUARD_END
}

int vcp_residual(vcp_ctx *c, int mon, double md[8])
{
  VCP_GUARD_BEGIN
  if (!c) return fail("null context");
  if (int rc_ = not_in_batch(c)) return rc_;
  int rc = 0;
  vcp::StageCtl ctl{1, 0, 0.0, 0.0, mon ? 1 : 0};
  if ((rc = residual_pass(c, ctl))) return rc;
  double sc[vcp::kScalCount] = {};
  HIP_TRY(hipMemcpyAsync(sc, c->dc.scal, kScalRead * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if ((rc = check_flags(c))) return rc;
  if (md) fill_monitoring(c, md, mon != 0, sc);
  c->res_valid = true;
  return 0;
  VCP_GUARD_END
}

int vcp_residual_static(vcp_ctx *c, int mon, double md[8])
{
  VCP_GUARD_BEGIN
  if (!c) return fail("null context");
  if (int rc_ = not_in_batch(c)) return rc_;
  if (c->prm.sponge_layer && !c->damping_set)
    return fail("sponge layer damping state not set (vcp_set_damping_state)");
  // the face-slot buffer of the residual in the library's fixed device array
  const size_t need = (size_t) c->dc.nElem * 6 * 10 * c->dc.K * c->dc.K;
  void *scr = nullptr;
  size_t cap = 0;
  HIP_TRY(vcp::static_scratch(&scr, &cap));
  if (need * sizeof(double) > cap)
    return fail("vcp_residual_static: the face-slot buffer of this case (" + std::to_string(need * sizeof(double)) +
                    " bytes) exceeds the fixed scratch array (" + std::to_string(cap) + " bytes)", -7);
  vcp::DeviceCase d = c->dc;
  d.faceRes = static_cast<double *>(scr);
  // the array keeps the last call's contents: start from zero, as a new context does
  HIP_TRY(vcp::launch_fill(d.faceRes, need, 0.0, c->stream));
  const vcp::StageCtl ctl{1, 0, 0.0, 0.0, mon ? 1 : 0};
  HIP_TRY(vcp::launch_boundary(d, c->stream, mon != 0));
  HIP_TRY(vcp::launch_faces(d, c->stream));
  HIP_TRY(vcp::launch_volume(d, ctl, c->stream));
  double sc[vcp::kScalCount] = {};
  HIP_TRY(hipMemcpyAsync(sc, c->dc.scal, kScalRead * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  int rc = 0;
  if ((rc = check_flags(c))) return rc;
  if (md) fill_monitoring(c, md, mon != 0, sc);
  c->res_valid = true;
  return 0;
  VCP_GUARD_END
}

int vcp_get_geometry(vcp_ctx *c, double *vgeo, double *fgeo, double *econst)
{
  VCP_GUARD_BEGIN
  if (!c) return fail("null context");
  const vcp::DeviceCase &dc = c->dc;
  if (vgeo)
    HIP_TRY(hipMemcpy(vgeo, dc.vgeo, (size_t) dc.nElem * vcp::kVolGeo * dc.nInt * sizeof(double), hipMemcpyDeviceToHost));
  if (fgeo)
    HIP_TRY(hipMemcpy(fgeo, dc.fgeo, (size_t) dc.nElem * 3 * vcp::kFaceGeo * dc.nInt2D * sizeof(double),
                      hipMemcpyDeviceToHost));
  if (econst)
    HIP_TRY(hipMemcpy(econst, dc.econst, (size_t) dc.nElem * vcp::kElemConst * sizeof(double), hipMemcpyDeviceToHost));
  return 0;
  VCP_GUARD_END
}

int vcp_time_steps(vcp_ctx *c, double *dt_elem)
{
  VCP_GUARD_BEGIN
  if (!c) return fail("null context");
  if (int rc_ = not_in_batch(c)) return rc_;
  c->dt_ready = false;
  HIP_TRY(vcp::launch_time_step(c->dc, c->stream));
  if (dt_elem)
    HIP_TRY(hipMemcpyAsync(dt_elem, c->dc.dtElem, c->dc.nElem * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  return check_flags(c);
  VCP_GUARD_END
}

int vcp_sync_step(vcp_ctx *c, double md[8])
{
  VCP_GUARD_BEGIN
  if (!c) return fail("null context");
  if (int rc_ = not_in_batch(c)) return rc_;
  double dummy[8];
  if (!md) md = dummy;
  int rc = 0;
  if (!c->res_valid && (rc = vcp_residual(c, 1, md))) return rc;
  const double dtSync = c->dt_sync;
  double elapsed = 0.0;
  bool reached = false;
  int nSteps = 0;
  while (!reached) {
    ++c->time_step; ++nSteps;
    double dt = 0.0;
    if (c->dt_ready) {
      // computed, with its error flags checked, at the end of the previous call
      dt = c->host[0];
      c->dt_ready = false;
    } else {
      // time step and error flags back in one round trip
      HIP_TRY(vcp::launch_time_step(c->dc, c->stream));
      HIP_TRY(hipMemcpyAsync(c->host, c->dc.scal, sizeof(double), hipMemcpyDeviceToHost, c->stream));
      HIP_TRY(hipMemcpyAsync(c->host + kHostFlags, c->dc.flags, sizeof(int), hipMemcpyDeviceToHost, c->stream));
      HIP_TRY(hipStreamSynchronize(c->stream));
      int flags = 0;
      std::memcpy(&flags, c->host + kHostFlags, sizeof(int));
      if ((rc = flag_error(flags))) return rc;
      dt = c->host[0];
    }
    if (elapsed + dt >= 0.99 * dtSync) { dt = dtSync - elapsed; reached = true; }
    elapsed += dt;
    std::swap(c->dc.U, c->dc.U0);   // the step's state becomes U0, stage 0 writes U
    if ((rc = c->use_graph ? rk_graph(c, reached, dt) : rk_stages(c, reached, &dt))) return rc;
    if (reached && c->dc.avg) {
      // running averages of the new state (SolverClass.cpp:2415-2421)
      ++c->n_avg;
      HIP_TRY(vcp::launch_average(c->dc, 1.0 / c->n_avg, c->stream));
    }
    if (reached) {
      // the next call's time step with the monitoring maximum and the error
      // flags in one round trip
      HIP_TRY(hipMemcpyAsync(c->host, c->dc.scal, kScalRead * sizeof(double), hipMemcpyDeviceToHost, c->stream));
      HIP_TRY(hipMemcpyAsync(c->host + kHostFlags, c->dc.flags, sizeof(int), hipMemcpyDeviceToHost, c->stream));
      HIP_TRY(hipStreamSynchronize(c->stream));
      if (c->kt.on && (rc = collect_timing(c))) return rc;
      int flags = 0;
      std::memcpy(&flags, c->host + kHostFlags, sizeof(int));
      // a failure of the next step's time step is reported by the next call
      if ((rc = flag_error(flags & ~vcp::kFlagNegativeA2))) return rc;
      c->dt_ready = !(flags & vcp::kFlagNegativeA2);
      fill_monitoring(c, md, true, c->host);
      if (c->dc.avg) {
        // averaged force coefficients from the step's last monitoring data
        // (SolverClass.cpp:3858-3869)
        const double factNew = 1.0 / c->n_avg, factOld = 1.0 - factNew;
        for (int i = 0; i < 6; ++i) c->cf_avg[i] = factOld * c->cf_avg[i] + factNew * md[i];
      }
    }
  }
  c->res_valid = true;
  return nSteps;
  VCP_GUARD_END
}

int vcp_get_averages(vcp_ctx *c, double *avg, int *n_steps)
{
  VCP_GUARD_BEGIN
  if (!c || !avg || !n_steps) return fail("null argument");
  if (int rc_ = vcp::batch_quiesce(c)) return rc_;
  if (!c->dc.avg) return fail("the case was created without compute_average");
  const int nD = c->dc.nDOF, na = c->dc.nAvg;
  std::vector<double> buf((size_t) c->dc.nElem * na * nD);
  HIP_TRY(hipMemcpyAsync(buf.data(), c->dc.avg, buf.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  for (int e = 0; e < c->dc.nElem; ++e)
    for (int d = 0; d < nD; ++d) {
      double *o = avg + ((size_t) e * nD + d) * 12;
      for (int i = 0; i < na; ++i) o[i] = buf[((size_t) e * na + i) * nD + d];
      if (na == 11) o[11] = 0.0;   // no SGS model: the averaged eddy viscosity is 0
    }
  *n_steps = c->n_avg;
  return 0;
  VCP_GUARD_END
}

int vcp_set_averages(vcp_ctx *c, const double *avg, int n_steps)
{
  VCP_GUARD_BEGIN
  if (!c || !avg) return fail("null argument");
  if (int rc_ = not_in_batch(c)) return rc_;
  if (!c->dc.avg) return fail("the case was created without compute_average");
  if (n_steps < 0) return fail("negative number of averaged steps");
  const int nD = c->dc.nDOF, na = c->dc.nAvg;
  std::vector<double> buf((size_t) c->dc.nElem * na * nD);
  for (int e = 0; e < c->dc.nElem; ++e)
    for (int d = 0; d < nD; ++d)
      for (int i = 0; i < na; ++i) buf[((size_t) e * na + i) * nD + d] = avg[((size_t) e * nD + d) * 12 + i];
  HIP_TRY(hipMemcpyAsync(c->dc.avg, buf.data(), buf.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->n_avg = n_steps;
  return 0;
  VCP_GUARD_END
}

int vcp_set_kernel_timing(vcp_ctx *c, int enable)
{
  VCP_GUARD_BEGIN
  if (!c) return fail("null context");
  if (int rc_ = not_in_batch(c)) return rc_;
  vcp_ctx::KernelTiming &kt = c->kt;
  kt.on = enable != 0;
  kt.used = 0;
  if (kt.on) {
    kt.face_ms = kt.vol_ms = 0.0;
    kt.launches = 0;
  }
  return 0;
  VCP_GUARD_END
}

int vcp_get_kernel_timing(vcp_ctx *c, double ms[2], int *launches)
{
  VCP_GUARD_BEGIN
  if (!c || !ms || !launches) return fail("null argument");
  const vcp_ctx::KernelTiming &kt = c->kt;
  *launches = kt.launches;
  ms[0] = kt.launches ? kt.face_ms / kt.launches : 0.0;
  ms[1] = kt.launches ? kt.vol_ms / kt.launches : 0.0;
  return 0;
  VCP_GUARD_END
}

int vcp_get_force_averages(vcp_ctx *c, double cf[6])
{
  VCP_GUARD_BEGIN
  if (!c || !cf) return fail("null argument");
  if (!c->dc.avg) return fail("the case was created without compute_average");
  for (int i = 0; i < 6; ++i) cf[i] = c->cf_avg[i];
  return 0;
  VCP_GUARD_END
}

int vcp_set_force_averages(vcp_ctx *c, const double cf[6])
{
  VCP_GUARD_BEGIN
  if (!c || !cf) return fail("null argument");
  if (!c->dc.avg) return fail("the case was created without compute_average");
  for (int i = 0; i < 6; ++i) c->cf_avg[i] = cf[i];
  return 0;
  VCP_GUARD_END
}

int vcp_set_graph(vcp_ctx *c, int enable)
{
  VCP_GUARD_BEGIN
  if (!c) return fail("null context");
  if (int rc_ = not_in_batch(c)) return rc_;
  c->use_graph = enable != 0;
  return 0;
  VCP_GUARD_END
}

int vcp_set_tile(vcp_ctx *c, int nt)
{
  VCP_GUARD_BEGIN
  if (!c) return fail("null context");
  if (int rc_ = not_in_batch(c)) return rc_;
  if (nt != 0) {
    int t[4];
    const int n = vcp::volume_tiles(c->dc.K, c->dc.M, t);
    bool ok = false;
    for (int i = 0; i < n; ++i) ok = ok || t[i] == nt;
    if (!ok) return fail("volume workgroup size not available for this (nDOFs1D, nInt1D)");
  }
  c->dc.nt = nt;
  c->drop_graphs();
  return 0;
  VCP_GUARD_END
}

int vcp_device_bytes(const vcp_ctx *c, size_t *bytes)
{
  VCP_GUARD_BEGIN
  if (!c || !bytes) return fail("null argument");
  *bytes = c->bytes;
  return 0;
  VCP_GUARD_END
}

int vcp_set_precision(vcp_ctx *c, int mode)
{
  VCP_GUARD_BEGIN
  if (!c) return fail("null context");
  if (int rc_ = not_in_batch(c)) return rc_;
  if (mode != VCP_PRECISION_FP64 && mode != VCP_PRECISION_MIXED) return fail("unknown precision mode");
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipStreamSynchronize(c->side));
  c->dc.mixed = mode == VCP_PRECISION_MIXED ? 1 : 0;
  c->drop_graphs();
  return 0;
  VCP_GUARD_END
}

int vcp_volume_tiles(int n_dofs_1d, int n_int_1d, int out[4])
{
  VCP_GUARD_BEGIN
  if (!out) return fail("null argument");
  return vcp::volume_tiles(n_dofs_1d, n_int_1d, out);
  VCP_GUARD_END
}

int vcp_cg_iterations(vcp_ctx *c, int *iters)
{
  VCP_GUARD_BEGIN
  if (!c || !iters) return fail("null argument");
  if (int rc_ = vcp::batch_quiesce(c)) return rc_;
  HIP_TRY(hipMemcpyAsync(iters, c->dc.cgIter, c->dc.nElem * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return 0;
  VCP_GUARD_END
}

int vcp_profile_residual(vcp_ctx *c, int reps, double ms[2])
{
  VCP_GUARD_BEGIN
  if (!c || !ms || reps <= 0) return fail("invalid argument");
  if (int rc_ = not_in_batch(c)) return rc_;
  vcp::EventSet es;
  HIP_TRY(es.create(3 * (size_t) reps));
  const std::vector<hipEvent_t> &ev = es.ev;
  vcp::StageCtl ctl{1, 0, 0.0, 0.0, 0};
  for (int r = 0; r < reps; ++r) {
    HIP_TRY(hipEventRecord(ev[3 * r], c->stream));
    HIP_TRY(vcp::launch_boundary(c->dc, c->stream));
    HIP_TRY(vcp::launch_faces(c->dc, c->stream));
    HIP_TRY(hipEventRecord(ev[3 * r + 1], c->stream));
    HIP_TRY(vcp::launch_volume(c->dc, ctl, c->stream));
    HIP_TRY(hipEventRecord(ev[3 * r + 2], c->stream));
  }
  HIP_TRY(hipStreamSynchronize(c->stream));
  double f = 0.0, v = 0.0;
  for (int r = 0; r < reps; ++r) {
    float a = 0.f, b = 0.f;
    HIP_TRY(hipEventElapsedTime(&a, ev[3 * r], ev[3 * r + 1]));
    HIP_TRY(hipEventElapsedTime(&b, ev[3 * r + 1], ev[3 * r + 2]));
    f += a; v += b;
  }
  ms[0] = f / reps; ms[1] = v / reps;
  c->res_valid = true;
  return check_flags(c);
  VCP_GUARD_END
}

}  // extern "C"
