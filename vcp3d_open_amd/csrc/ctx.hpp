// Internal state behind the C ABI: one case context (vcp_ctx) and the error
// plumbing shared by the translation units of the library (capi.cpp,
// batch.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <exception>
#include <memory>
#include <string>
#include <vector>

#include "kernels.hpp"
#include "setup.hpp"
#include "vcp3d_amd.h"

namespace vcp {

// Records msg as the text of vcp_last_error() and returns code.
int fail(const std::string &msg, int code = -1);

}  // namespace vcp

#define HIP_TRY(expr)                                                              \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess)                                                          \
      return vcp::fail(std::string(#expr) + ": " + hipGetErrorString(e_), -3);     \
  } while (0)

// Exported functions catch every C++ exception (std::bad_alloc, a failed
// std::thread, a singular table matrix in the set-up) and turn it into an
// error code, so nothing unwinds through the C ABI.
#define VCP_GUARD_BEGIN try {
#define VCP_GUARD_END                                                              \
  }                                                                                \
  catch (const std::exception &ex_) { return vcp::fail(std::string("exception: ") + ex_.what(), -2); } \
  catch (...) { return vcp::fail("unknown exception", -2); }

struct vcp_ctx {
  vcp::CaseParams prm;
  vcp::Tables tab;
  vcp::RefConditions rc{};
  vcp::DeviceCase dc;
  vcp::Sponge sponge;
  hipStream_t stream = nullptr;
  // side stream for the next call's time step, which overlaps the last stage
  hipStream_t side = nullptr;
  hipEvent_t evState = nullptr, evDt = nullptr;
  // pinned host words the per-step scalars come back through: [0, kScalRead)
  // the device scalars (time step, maximum Mach^2, ..., wall forces),
  // [kHostFlags] the error flags (as int)
  double *host = nullptr;
  std::vector<void *> allocs;
  size_t bytes = 0;               // device bytes requested by alloc()
  bool res_valid = false;
  bool dt_ready = false;          // host[0] holds the time step of the current state
  bool damping_set = false;
  // non-dimensional synchronization time step, rounded once at creation: the
  // host loop of vcp_sync_step and a batch's device control both clip with
  // this value (formed inside the clipping expression, the product could be
  // fused into the subtraction and round differently)
  double dt_sync = 0.0;
  int time_step = 0;
  int n_avg = 0;                  // synchronization steps in the running averages
  // running averages of the six force coefficients of monitoring[0..5]
  // (SolverClass::mCfxAve .. mCfzVisAve)
  double cf_avg[6] = {};
  // RK steps replayed as captured HIP graphs (vcp_set_graph), one per
  // [state buffer that is current][last step of the call], captured on first use
  bool use_graph = false;
  double *U_first = nullptr;
  hipGraphExec_t graph[2][2] = {};
  // the batch that steps this case (vcp_batch_add), or null
  vcp_batch *batch = nullptr;
  // launch timing inside vcp_sync_step (vcp_set_kernel_timing): three events
  // per RK stage (before the face launch, between it and the volume launch,
  // after the volume launch), read back after each call's last synchronisation
  struct KernelTiming {
    bool on = false;
    std::vector<hipEvent_t> ev;
    size_t used = 0;
    double face_ms = 0.0, vol_ms = 0.0;
    int launches = 0;
  } kt;

  template <class T>
  hipError_t alloc(T **p, size_t n)
  {
    void *v = nullptr;
    const size_t nb = n * sizeof(T) > 0 ? n * sizeof(T) : 16;
    hipError_t e = hipMalloc(&v, nb);
    if (e == hipSuccess) { allocs.push_back(v); bytes += nb; *p = static_cast<T *>(v); }
    return e;
  }
  void drop_graphs()
  {
    for (auto &row : graph)
      for (hipGraphExec_t &g : row)
        if (g) { (void) hipGraphExecDestroy(g); g = nullptr; }
  }
  ~vcp_ctx()
  {
    drop_graphs();
    for (hipEvent_t e : kt.ev) (void) hipEventDestroy(e);
    for (void *p : allocs) (void) hipFree(p);
    if (host) (void) hipHostFree(host);
    if (evState) (void) hipEventDestroy(evState);
    if (evDt) (void) hipEventDestroy(evDt);
    if (side) (void) hipStreamDestroy(side);
    if (stream) (void) hipStreamDestroy(stream);
  }
};

namespace vcp {

// HIP events destroyed with the holder on every exit path.
struct EventSet {
  std::vector<hipEvent_t> ev;
  hipError_t create(size_t n)
  {
    ev.assign(n, nullptr);
    for (hipEvent_t &e : ev) {
      const hipError_t r = hipEventCreate(&e);
      if (r != hipSuccess) { e = nullptr; return r; }
    }
    return hipSuccess;
  }
  ~EventSet()
  {
    for (hipEvent_t e : ev)
      if (e) (void) hipEventDestroy(e);
  }
};

// open boundaries + faces + volume/mass matrix of one case, with the given
// stage control, on stream s
int residual_pass(vcp_ctx *c, const StageCtl &ctl, hipStream_t s);
// waits for the batch of c (if any) to finish its queued work
int batch_quiesce(vcp_ctx *c);

}  // namespace vcp
