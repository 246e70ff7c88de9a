// Device-side data of one case and the kernel launchers (gfx950, FP64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace vcp {

// Offsets (in doubles) of the 1D tables in the device table block.
struct TableOffsets {
  int A, D, AT, DT, LD, DLD, fmin, fmax, dfmin, dfmax, w3, w2, bc, P1, dLq, Lq, total;
};

// PC-NSCBC parameters of the block faces (InputParamClass "NSCBC tuning
// parameters" and the prescribed data, non-dimensional).
struct NscbcParams {
  double outlet_len, outlet_sigma, outlet_beta_l, outlet_beta_t;
  double inlet_len, inlet_sigma, dell0;
  int outlet_one_minus_beta, inlet_incoming;
  int bc[6];
  int begin[7];                 // boundary faces of block face f: [begin[f], begin[f+1])
  double data[6][5];
  double invArea[6];
};

// Per-face geometry slot layout, [nElem][3][kFaceGeo][nInt2D]:
//   0..2 unit normal (L -> R), 3 area factor, 4..12 inverse metric of the
//   left (max-face) element, 13..21 inverse metric of the right (min-face)
//   element.
constexpr int kFaceGeo = 22;
// Per-element volume geometry, [nElem][kVolGeo][nInt]: J, dr/dx .. dt/dz.
constexpr int kVolGeo = 10;
// Per-element constants, [nElem][kElemConst]: len I,J,K, volume, aveNorm[9],
// the LES length scale (slot kElemLenLES).
constexpr int kElemConst = 16, kElemLenLES = 13;
// Per-element metric at the nodal solution DOFs (SGS model only),
// [nElem][kNodalGeo][nDOF]: J, dr/dx .. dt/dz.
constexpr int kNodalGeo = 10;
// Subgrid scale models (DeviceCase::sgs).
enum : int { kSgsNone = 0, kSgsWale = 1, kSgsVreman = 2 };

// Device status flags.
enum : int { kFlagCGNotConverged = 1, kFlagNegativeA2 = 2, kFlagNonFinite = 4 };

struct DeviceCase {
  int nElem = 0, K = 0, M = 0, nDOF = 0, nInt = 0, nInt2D = 0, P = 0;
  int bug_compat = 1;
  int nt = 0;                     // volume workgroup size (0: the default for (K, M))
  int mixed = 0;                  // 1: FP32 contractions and solve (vcp_set_precision)
  int sgs = kSgsNone;             // subgrid scale model
  int nAvg = 11;                  // averaged quantities per DOF (12 with an SGS model)
  int fbodyOn = 0;                // body force per unit volume set
  double fbody[3] = {0, 0, 0};    // its non-dimensional value
  double *bfBuf = nullptr;        // [nElem][4][nInt] its point sources (fbodyOn)
  double mu = 0, theta_sym = 1, rel_penalty = 1;
  double cfl = 1, u_ref = 1, p_ref = 1, rho_ref = 1;
  TableOffsets to{};
  double opA[128] = {}, opD[128] = {};   // host copies of the M x K operators
  double opF[4][16] = {};                // phi(-1), phi(+1), phi'(-1), phi'(+1)
  double opW[16] = {};                   // 1D Gauss weights
  double opC0[16] = {};                  // phi_q(0), the basis at the element centre
  double *tables = nullptr;
  double *vgeo = nullptr, *fgeo = nullptr, *econst = nullptr;
  double *vgeoN = nullptr;        // [nElem][kNodalGeo][nDOF] (SGS model only)
  // eddy-viscosity maxima of the last residual (SGS model only):
  // [0, nElem) volume points per element, [nElem, 4 nElem) internal-face
  // points per (element, direction), [4 nElem, 4 nElem + nBFace) boundary faces
  double *eddyPart = nullptr;
  int *nbr = nullptr;
  double *U = nullptr, *U0 = nullptr, *R = nullptr;
  double *faceRes = nullptr;      // [nElem][6][2][5][K][K] face-slot coefficient planes
  // device scalars: [0] minimum time step, [1] maximum Mach^2, [kScalDt] the
  // time step of the RK step being taken (written by the host before its
  // stages, so that the launches of a step do not depend on its value),
  // [kScalForce .. +6) the wall force sums of the last monitored residual
  double *scal = nullptr;
  int *flags = nullptr;
  int *cgIter = nullptr;          // [nElem]
  double *dtElem = nullptr;       // [nElem]
  double *mach2Elem = nullptr;    // [nElem] maximum Mach^2 of the last monitored residual
  // open boundaries
  int nBFace = 0;
  int *bface = nullptr;           // [nBFace][2]: element, block face
  double *bgeo = nullptr;         // [nBFace][13][nInt2D]: the face's normal, area factor, inverse metric
  double *machPart = nullptr;     // [nBFace] weighted face-normal Mach number
  double *forcePart = nullptr;    // [nBFace][6] wall forces (pressure xyz, viscous xyz)
  double *avg = nullptr;          // [nElem][nAvg][nDOF] running averages (null: not computed)
  int monitor_forces = 0;
  NscbcParams nscbc{};
  // sponge layer / characteristic matching layer
  int nSponge = 0, cml = 0;
  int *spIdx = nullptr;           // [nElem] slot or -1
  double *spCoef = nullptr;       // [nSponge][8]
  double *dampFn = nullptr;       // [nSponge][3][nInt]
  double *dampState = nullptr;    // [nSponge][5][nInt]
  double *srcBuf = nullptr;       // [nSponge][5][nInt] source terms of the current residual
};

// Device scalars of a case (DeviceCase::scal): [kScalDt], [kScalForce ..
// +6) as above; [kScalEddy] the maximum eddy viscosity of the last monitored
// residual (SGS model); the per-case control of a batch: [kScalElapsed] time elapsed
// in the current synchronization step, [kScalSyncDone] synchronization steps
// completed, [kScalDtSync] the synchronization time step (non-dimensional).
constexpr int kScalDt = 2, kScalForce = 3, kScalElapsed = 9, kScalSyncDone = 10, kScalDtSync = 12,
              kScalEddy = 13, kScalCount = 16;
// Device words of a case (DeviceCase::flags): [kWordFlags] error flags,
// [kWordActive] 1 while the case is stepped (always outside a batch; a batch
// clears it for a case whose budget is spent or whose error flags are set, and
// its kernels then skip the case), [kWordStepsLeft] RK steps the batch still
// takes for it, [kWordStepsDone] RK steps it took.
constexpr int kWordFlags = 0, kWordActive = 1, kWordStepsLeft = 2, kWordStepsDone = 3, kWordCount = 4;

// Per-case device pointers and scalars: the case part of the face, volume,
// time-step and average kernels' arguments, and the entries of a batch's
// device case table.
struct CaseDev {
  double *U, *U0, *R, *faceRes;
  const double *tables, *vgeo, *fgeo, *econst, *vgeoN;
  const int *nbr;
  double *scal;
  int *flags, *cgIter;
  double *dtElem, *mach2Elem;
  const int *spIdx;               // sponge slot per element (null: no sponge)
  const double *spCoef, *dampFn, *dampState;
  double *srcBuf;
  double *eddyPart;               // eddy-viscosity maxima (null: no SGS model)
  double mu, tauFact, thetaHalf, cfl, u_ref, p_ref, rho_ref, muDim;
  double *bfBuf;                  // body-force point sources [nElem][4][nInt] (null: none)
  double fbody[3];                // non-dimensional body force
  int nElem, cml, sgs;
};

// Stage control for the fused volume/mass-matrix kernel.
struct StageCtl {
  int store_res;      // write R = M^-1 res
  int update;         // U = (1-a) U + a U0 - b dt R in place, dt = scal[kScalDt]
  double a, b;
  int monitor;        // Mach^2 max
};

CaseDev case_dev(const DeviceCase &dc);

// Work item of a batched launch: (index into the group's case table, element).
struct WorkItem {
  int c, e;
};

// One shape group of a batch: cases with the same (K, M) and metric-indexing
// mode, run by one launch per kernel over a work list of (case, element)
// items. proto is a member case: the 1D operators and table offsets, which
// depend on (K, M) only.
struct BatchGroup {
  const DeviceCase *proto = nullptr;
  const CaseDev *cases = nullptr;   // device [nCase]
  const WorkItem *work = nullptr;   // device [nWork], ordered by case then element
  int nCase = 0, nWork = 0;
  int nt = 0;                       // volume workgroup size, 0: the default
};

bool supported_pair(int K, int M);
// Volume workgroup sizes available for (K, M); the first is the default.
int volume_tiles(int K, int M, int out[4]);
int default_volume_tile(int K, int M);
hipError_t launch_faces(const DeviceCase &dc, hipStream_t s);
// boundary-averaged Mach number partial sums, then the PC-NSCBC faces
// with monitor, also the wall forces into scal[kScalForce ..]
hipError_t launch_boundary(const DeviceCase &dc, hipStream_t s, bool monitor = false);
hipError_t launch_volume(const DeviceCase &dc, const StageCtl &ctl, hipStream_t s);
hipError_t launch_time_step(const DeviceCase &dc, hipStream_t s);
// running averages avg = (1 - factNew) avg + factNew q of the current state
hipError_t launch_average(const DeviceCase &dc, double factNew, hipStream_t s);
// sponge target (dampState) of every sponge element from the running
// averages of the primitive variables (ElementClass::InitializeDampingState)
hipError_t launch_damping_from_averages(const DeviceCase &dc, hipStream_t s);
// scal[kScalDt] = dt, ordered on the stream
hipError_t launch_set_dt(const DeviceCase &dc, double dt, hipStream_t s);
// p[0, n) = v on stream s
hipError_t launch_fill(double *p, size_t n, double v, hipStream_t s);
// The library's fixed device scratch array (64 MiB per device, held until the
// process ends): its address on the current device, looked up once per device
// with hipGetSymbolAddress, and its size. No allocation.
constexpr size_t kStaticScratchBytes = size_t(64) << 20;
constexpr int kMaxDevices = 64;
hipError_t static_scratch(void **ptr, size_t *bytes);
// U = U0 - dt R: with dt given, the kernel also stores it to scal[kScalDt];
// with dt null it reads scal[kScalDt]
hipError_t launch_stage0(const DeviceCase &dc, const double *dt, hipStream_t s);

// Batched launches over the work items [w0, w0 + nw) of a group (cases whose
// kWordActive word is 0 are skipped).
hipError_t launch_batch_faces(const BatchGroup &g, int w0, int nw, hipStream_t s);
hipError_t launch_batch_volume(const BatchGroup &g, const StageCtl &ctl, int w0, int nw, hipStream_t s);
// per-element time steps, then per case: the minimum, the synchronization-step
// clipping into scal[kScalDt], and the step counters; a case with error flags
// or no steps left is deactivated instead
hipError_t launch_batch_time_step(const BatchGroup &g, hipStream_t s);
// U0 <- U, U <- U - dt R with each case's scal[kScalDt]
hipError_t launch_batch_stage0(const BatchGroup &g, hipStream_t s);
// per case of the group: out[c][kStatusWords] = active, RK steps done, RK
// steps left, error flags, synchronization steps done
constexpr int kStatusWords = 5;
hipError_t launch_batch_status(const BatchGroup &g, int *out, hipStream_t s);

}  // namespace vcp
