// Batched stepping of many independent cases (vcp_batch_*).
//
// The reference runs one case per SolverClass (SolverClass.hpp:40-124) and
// dispatches each tensor-product call on (nDOFs1D, nInt1D) at run time
// (TensorProductSolAndGradVolume.cpp:27-56). Here a batch groups its cases by
// (K, M, metric-indexing mode, precision); each group is one work list of (case,
// element) items over a device table of the cases' buffers, so every kernel
// of an RK stage is one launch per group whatever the number of cases. The
// per-case time step, the synchronization-step clipping of
// ExecuteTimeSynchronizationStep (SolverClass.cpp:2356-2365) and the step
// budget are kept on the device, so an RK step of the whole batch needs no
// host round trip and can be replayed as one captured HIP graph. A case whose
// error flags are set (the reference would Terminate: CG_MassMatrix.cpp:
// 398-400, ElementClass.cpp:1631-1633) or whose step budget is spent is
// deactivated on the device and skipped; the others go on.
#include "ctx.hpp"

#include <algorithm>
#include <climits>
#include <cstring>

using vcp::fail;

struct vcp_batch {
  struct Group {
    int K = 0, M = 0, bug = 0, mixed = 0, sgs = 0, nt = 0;   // sgs: any SGS model (time-step kernel)
    std::vector<vcp_ctx *> cases;      // case table order
    std::vector<int> first;            // first work item of each case, + total
    vcp::CaseDev *dCases = nullptr;
    size_t capCases = 0;
    vcp::WorkItem *dWork = nullptr;
    size_t capWork = 0;
    int *dStatus = nullptr;
    int nWork = 0;
    int initFrom = -1;                 // cases [initFrom, end) need their first residual

    vcp::BatchGroup view() const
    {
      vcp::BatchGroup g;
      g.proto = cases.empty() ? nullptr : &cases[0]->dc;
      g.cases = dCases; g.work = dWork;
      g.nCase = (int) cases.size(); g.nWork = nWork; g.nt = nt;
      return g;
    }
    void release()
    {
      if (dCases) (void) hipFree(dCases);
      if (dWork) (void) hipFree(dWork);
      if (dStatus) (void) hipFree(dStatus);
      dCases = nullptr; dWork = nullptr; dStatus = nullptr;
      capCases = capWork = 0;
    }
  };
  hipStream_t stream = nullptr;
  std::vector<Group> groups;
  bool dirty = false;                  // device tables out of date
  bool use_graph = false;
  hipGraphExec_t graph = nullptr;      // one RK step of the whole batch

  void drop_graph()
  {
    if (graph) (void) hipGraphExecDestroy(graph);
    graph = nullptr;
  }
  ~vcp_batch()
  {
    if (stream) (void) hipStreamSynchronize(stream);
    drop_graph();
    for (Group &g : groups) {
      for (vcp_ctx *c : g.cases) c->batch = nullptr;
      g.release();
    }
    if (stream) (void) hipStreamDestroy(stream);
  }
};

namespace {

// SSP-RK3 stage coefficients (Shu-Osher form, SolverClass.cpp:2192-2194)
const vcp::StageCtl kStage1{0, 1, 3.0 / 4.0, 1.0 / 4.0, 0};
const vcp::StageCtl kStage2{0, 1, 1.0 / 3.0, 2.0 / 3.0, 0};
const vcp::StageCtl kStage3{1, 0, 0.0, 0.0, 0};
const vcp::StageCtl kResidualOnly{1, 0, 0.0, 0.0, 0};

template <class T>
hipError_t grow(T **p, size_t &cap, size_t n)
{
  if (n <= cap && *p) return hipSuccess;
  if (*p) (void) hipFree(*p);
  *p = nullptr;
  size_t c = std::max<size_t>(n, 2 * cap);
  if (c == 0) c = 1;
  const hipError_t e = hipMalloc((void **) p, c * sizeof(T));
  cap = e == hipSuccess ? c : 0;
  return e;
}

// Device case tables and work lists of every group from the host lists.
int upload(vcp_batch *b)
{
  if (!b->dirty) return 0;
  HIP_TRY(hipStreamSynchronize(b->stream));
  b->drop_graph();
  for (vcp_batch::Group &g : b->groups) {
    std::vector<vcp::CaseDev> tab;
    std::vector<vcp::WorkItem> work;
    g.first.assign(1, 0);
    for (size_t ci = 0; ci < g.cases.size(); ++ci) {
      const vcp::DeviceCase &dc = g.cases[ci]->dc;
      tab.push_back(vcp::case_dev(dc));
      for (int e = 0; e < dc.nElem; ++e) work.push_back({(int) ci, e});
      g.first.push_back((int) work.size());
    }
    g.nWork = (int) work.size();
    if (g.cases.empty()) continue;
    const size_t capBefore = g.capCases;
    HIP_TRY(grow(&g.dCases, g.capCases, tab.size()));
    if (g.capCases != capBefore || !g.dStatus) {
      if (g.dStatus) (void) hipFree(g.dStatus);
      g.dStatus = nullptr;
      HIP_TRY(hipMalloc((void **) &g.dStatus, g.capCases * vcp::kStatusWords * sizeof(int)));
    }
    HIP_TRY(grow(&g.dWork, g.capWork, work.size()));
    HIP_TRY(hipMemcpy(g.dCases, tab.data(), tab.size() * sizeof(vcp::CaseDev), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(g.dWork, work.data(), work.size() * sizeof(vcp::WorkItem), hipMemcpyHostToDevice));
  }
  b->dirty = false;
  return 0;
}

// The residual of the cases that joined since the last step (their first
// RK stage reuses it, as after SolverClass::Residual in the reference's main).
int init_pass(vcp_batch *b)
{
  for (vcp_batch::Group &g : b->groups) {
    if (g.initFrom < 0) continue;
    const int w0 = g.first[g.initFrom], nw = g.nWork - w0;
    const vcp::BatchGroup v = g.view();
    for (size_t ci = g.initFrom; ci < g.cases.size(); ++ci)
      HIP_TRY(vcp::launch_boundary(g.cases[ci]->dc, b->stream, false));
    HIP_TRY(vcp::launch_batch_faces(v, w0, nw, b->stream));
    HIP_TRY(vcp::launch_batch_volume(v, kResidualOnly, w0, nw, b->stream));
    g.initFrom = -1;
  }
  return 0;
}

// One SSP-RK3 step of every active case: per-element time steps and the
// per-case clipping, the first stage, then the three residual stages.
int rk_step(vcp_batch *b)
{
  for (const vcp_batch::Group &g : b->groups) HIP_TRY(vcp::launch_batch_time_step(g.view(), b->stream));
  for (const vcp_batch::Group &g : b->groups) HIP_TRY(vcp::launch_batch_stage0(g.view(), b->stream));
  for (const vcp::StageCtl *ctl : {&kStage1, &kStage2, &kStage3})
    for (const vcp_batch::Group &g : b->groups) {
      if (g.cases.empty()) continue;
      const vcp::BatchGroup v = g.view();
      for (vcp_ctx *c : g.cases) HIP_TRY(vcp::launch_boundary(c->dc, b->stream, false));
      HIP_TRY(vcp::launch_batch_faces(v, 0, v.nWork, b->stream));
      HIP_TRY(vcp::launch_batch_volume(v, *ctl, 0, v.nWork, b->stream));
    }
  return 0;
}

int rk_step_graph(vcp_batch *b)
{
  if (!b->graph) {
    hipGraph_t g = nullptr;
    HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    const int rc = rk_step(b);
    const hipError_t he = hipStreamEndCapture(b->stream, &g);
    if (rc || he != hipSuccess) {
      if (g) (void) hipGraphDestroy(g);
      return rc ? rc : fail(std::string("hipStreamEndCapture: ") + hipGetErrorString(he), -3);
    }
    const hipError_t hi = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
    (void) hipGraphDestroy(g);
    if (hi != hipSuccess) {
      b->graph = nullptr;
      return fail(std::string("hipGraphInstantiate: ") + hipGetErrorString(hi), -3);
    }
  }
  HIP_TRY(hipGraphLaunch(b->graph, b->stream));
  return 0;
}

vcp_batch::Group *group_of(vcp_batch *b, const vcp_ctx *c, size_t *index)
{
  for (vcp_batch::Group &g : b->groups)
    for (size_t i = 0; i < g.cases.size(); ++i)
      if (g.cases[i] == c) {
        if (index) *index = i;
        return &g;
      }
  return nullptr;
}

}  // namespace

namespace vcp {

int batch_quiesce(vcp_ctx *c)
{
  if (c && c->batch) HIP_TRY(hipStreamSynchronize(c->batch->stream));
  return 0;
}

}  // namespace vcp

extern "C" {

int vcp_batch_create(vcp_batch **out)
{
  VCP_GUARD_BEGIN
  if (!out) return fail("null argument");
  *out = nullptr;
  std::unique_ptr<vcp_batch> b(new vcp_batch());
  HIP_TRY(hipStreamCreate(&b->stream));
  *out = b.release();
  return 0;
  VCP_GUARD_END
}

int vcp_batch_destroy(vcp_batch *b)
{
  VCP_GUARD_BEGIN
  if (!b) return 0;
  // the members are stepped alone from now on; the batch is freed whatever
  // the outcome, and the first failure is reported
  hipError_t he = hipStreamSynchronize(b->stream);
  for (vcp_batch::Group &g : b->groups)
    for (vcp_ctx *c : g.cases) {
      int words[vcp::kWordCount] = {};
      hipError_t e = hipMemcpy(words, c->dc.flags, sizeof(words), hipMemcpyDeviceToHost);
      if (e == hipSuccess) {
        c->time_step += words[vcp::kWordStepsDone];
        words[vcp::kWordActive] = 1;
        words[vcp::kWordStepsLeft] = 0;
        e = hipMemcpy(c->dc.flags, words, sizeof(words), hipMemcpyHostToDevice);
      }
      if (he == hipSuccess) he = e;
      c->batch = nullptr;
      c->res_valid = true;
      c->dt_ready = false;
    }
  delete b;
  if (he != hipSuccess) return fail(std::string("vcp_batch_destroy: ") + hipGetErrorString(he), -3);
  return 0;
  VCP_GUARD_END
}

int vcp_batch_add(vcp_batch *b, vcp_ctx *c, int n_rk_steps)
{
  VCP_GUARD_BEGIN
  if (!b || !c) return fail("null argument");
  if (c->batch) return fail("the context already belongs to a batch");
  if (c->prm.sponge_layer && !c->damping_set)
    return fail("sponge layer damping state not set (vcp_set_damping_state)");
  if (c->dc.avg) return fail("a case that computes running averages cannot join a batch");
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipStreamSynchronize(c->side));
  // control words: active, step budget, counters; the synchronization time step
  int words[vcp::kWordCount] = {};
  HIP_TRY(hipMemcpy(words, c->dc.flags, sizeof(words), hipMemcpyDeviceToHost));
  if (words[vcp::kWordFlags] != 0) return fail("the context has error flags set");
  words[vcp::kWordActive] = 1;
  words[vcp::kWordStepsLeft] = n_rk_steps > 0 ? n_rk_steps : INT_MAX;
  words[vcp::kWordStepsDone] = 0;
  const double sc[4] = {0.0, 0.0, 0.0, c->dt_sync};   // elapsed, sync done, -, dtSync
  static_assert(vcp::kScalSyncDone == vcp::kScalElapsed + 1 && vcp::kScalDtSync == vcp::kScalElapsed + 3,
                "control scalars layout");
  HIP_TRY(hipMemcpy(c->dc.flags, words, sizeof(words), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(c->dc.scal + vcp::kScalElapsed, sc, sizeof(sc), hipMemcpyHostToDevice));
  vcp_batch::Group *grp = nullptr;
  for (vcp_batch::Group &g : b->groups)
    if (g.K == c->dc.K && g.M == c->dc.M && g.bug == c->dc.bug_compat && g.mixed == c->dc.mixed
        && g.sgs == (c->dc.sgs != 0))
      grp = &g;
  if (!grp) {
    b->groups.emplace_back();
    grp = &b->groups.back();
    grp->K = c->dc.K; grp->M = c->dc.M; grp->bug = c->dc.bug_compat; grp->mixed = c->dc.mixed;
    grp->sgs = c->dc.sgs != 0;
  }
  if (grp->initFrom < 0) grp->initFrom = (int) grp->cases.size();
  grp->cases.push_back(c);
  c->batch = b;
  b->dirty = true;
  return 0;
  VCP_GUARD_END
}

int vcp_batch_remove(vcp_batch *b, vcp_ctx *c)
{
  VCP_GUARD_BEGIN
  if (!b || !c) return fail("null argument");
  size_t i = 0;
  vcp_batch::Group *g = group_of(b, c, &i);
  if (!g) return fail("the context does not belong to this batch");
  HIP_TRY(hipStreamSynchronize(b->stream));
  int words[vcp::kWordCount] = {};
  HIP_TRY(hipMemcpy(words, c->dc.flags, sizeof(words), hipMemcpyDeviceToHost));
  c->time_step += words[vcp::kWordStepsDone];
  words[vcp::kWordActive] = 1;     // stepped alone from now on
  words[vcp::kWordStepsLeft] = 0;
  HIP_TRY(hipMemcpy(c->dc.flags, words, sizeof(words), hipMemcpyHostToDevice));
  g->cases.erase(g->cases.begin() + (long) i);
  if (g->initFrom >= 0) {
    if ((int) i < g->initFrom) --g->initFrom;
    if (g->initFrom >= (int) g->cases.size()) g->initFrom = -1;
  }
  c->batch = nullptr;
  c->res_valid = true;     // the last stage of every RK step stores the residual
  c->dt_ready = false;
  b->dirty = true;
  return 0;
  VCP_GUARD_END
}

int vcp_batch_step(vcp_batch *b, int n_steps)
{
  VCP_GUARD_BEGIN
  if (!b || n_steps < 0) return fail("invalid argument");
  int rc = 0;
  if ((rc = upload(b))) return rc;
  if ((rc = init_pass(b))) return rc;
  for (int s = 0; s < n_steps; ++s)
    if ((rc = b->use_graph ? rk_step_graph(b) : rk_step(b))) return rc;
  return 0;
  VCP_GUARD_END
}

int vcp_batch_wait(vcp_batch *b)
{
  VCP_GUARD_BEGIN
  if (!b) return fail("null argument");
  HIP_TRY(hipStreamSynchronize(b->stream));
  return 0;
  VCP_GUARD_END
}

int vcp_batch_status(vcp_batch *b, vcp_ctx **ctxs, int *status, int cap)
{
  VCP_GUARD_BEGIN
  if (!b || (cap > 0 && (!ctxs || !status))) return fail("null argument");
  int rc = 0;
  if ((rc = upload(b))) return rc;
  int n = 0;
  for (vcp_batch::Group &g : b->groups)
    if (!g.cases.empty()) HIP_TRY(vcp::launch_batch_status(g.view(), g.dStatus, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  for (vcp_batch::Group &g : b->groups) {
    if (g.cases.empty()) continue;
    std::vector<int> st(g.cases.size() * vcp::kStatusWords);
    HIP_TRY(hipMemcpy(st.data(), g.dStatus, st.size() * sizeof(int), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < g.cases.size(); ++i, ++n)
      if (n < cap) {
        ctxs[n] = g.cases[i];
        std::memcpy(status + (size_t) n * vcp::kStatusWords, &st[i * vcp::kStatusWords],
                    vcp::kStatusWords * sizeof(int));
      }
  }
  return n;
  VCP_GUARD_END
}

int vcp_batch_set_graph(vcp_batch *b, int enable)
{
  VCP_GUARD_BEGIN
  if (!b) return fail("null argument");
  b->use_graph = enable != 0;
  return 0;
  VCP_GUARD_END
}

int vcp_batch_set_tile(vcp_batch *b, int K, int M, int nt)
{
  VCP_GUARD_BEGIN
  if (!b) return fail("null argument");
  if (nt != 0) {
    int t[4];
    const int n = vcp::volume_tiles(K, M, t);
    bool ok = false;
    for (int i = 0; i < n; ++i) ok = ok || t[i] == nt;
    if (!ok) return fail("volume workgroup size not available for this (nDOFs1D, nInt1D)");
  }
  HIP_TRY(hipStreamSynchronize(b->stream));
  for (vcp_batch::Group &g : b->groups)
    if (g.K == K && g.M == M) g.nt = nt;
  b->drop_graph();
  return 0;
  VCP_GUARD_END
}

int vcp_batch_groups(vcp_batch *b, int *info, int cap)
{
  VCP_GUARD_BEGIN
  if (!b || (cap > 0 && !info)) return fail("null argument");
  int rc = 0;
  if ((rc = upload(b))) return rc;
  int n = 0;
  for (const vcp_batch::Group &g : b->groups) {
    if (n < cap) {
      int *o = info + 7 * n;
      o[0] = g.K; o[1] = g.M; o[2] = g.bug; o[3] = (int) g.cases.size(); o[4] = g.nWork;
      o[5] = g.nt ? g.nt : vcp::default_volume_tile(g.K, g.M);
      o[6] = g.mixed;
    }
    ++n;
  }
  return n;
  VCP_GUARD_END
}

int vcp_batch_profile(vcp_batch *b, int reps, double *ms, int cap)
{
  VCP_GUARD_BEGIN
  if (!b || !ms || reps <= 0) return fail("invalid argument");
  int rc = 0;
  if ((rc = upload(b))) return rc;
  if ((rc = init_pass(b))) return rc;
  const int ng = (int) b->groups.size();
  if (cap < ng) return fail("ms holds fewer than two entries per group");
  // residual evaluations that store nothing (no R, no update): the face
  // slots, CG counters and flags are the only writes
  const vcp::StageCtl probe{0, 0, 0.0, 0.0, 0};
  vcp::EventSet es;
  HIP_TRY(es.create(3 * (size_t) reps * (size_t) std::max(ng, 1)));
  for (int r = 0; r < reps; ++r)
    for (int gi = 0; gi < ng; ++gi) {
      const vcp_batch::Group &g = b->groups[gi];
      const vcp::BatchGroup v = g.view();
      hipEvent_t *ev = &es.ev[3 * ((size_t) r * ng + gi)];
      HIP_TRY(hipEventRecord(ev[0], b->stream));
      for (vcp_ctx *c : g.cases) HIP_TRY(vcp::launch_boundary(c->dc, b->stream, false));
      HIP_TRY(vcp::launch_batch_faces(v, 0, v.nWork, b->stream));
      HIP_TRY(hipEventRecord(ev[1], b->stream));
      HIP_TRY(vcp::launch_batch_volume(v, probe, 0, v.nWork, b->stream));
      HIP_TRY(hipEventRecord(ev[2], b->stream));
    }
  HIP_TRY(hipStreamSynchronize(b->stream));
  for (int gi = 0; gi < ng; ++gi) {
    double f = 0.0, v = 0.0;
    for (int r = 0; r < reps; ++r) {
      const hipEvent_t *ev = &es.ev[3 * ((size_t) r * ng + gi)];
      float a = 0.f, c = 0.f;
      HIP_TRY(hipEventElapsedTime(&a, ev[0], ev[1]));
      HIP_TRY(hipEventElapsedTime(&c, ev[1], ev[2]));
      f += a; v += c;
    }
    ms[2 * gi] = f / reps;
    ms[2 * gi + 1] = v / reps;
  }
  return ng;
  VCP_GUARD_END
}

}  // extern "C"
