// gfx950 kernels of the VCP3D explicit DG residual and SSP-RK3 stage (FP64).
//
//   face_kernel          interior (incl. periodic) faces: traces of both
//                        elements, Ismail-Roe + viscous + penalty +
//                        symmetrising fluxes, lifted to both elements into
//                        per-element face slots (no atomics)
//   volume_kernel        volume residual, gather of the six face slots,
//                        per-element entropy mass-matrix PCG, optional
//                        in-place RK stage update
//   mach_kernel          boundary-averaged Mach number partial sums (NSCBC)
//   nscbc_kernel         PC-NSCBC boundary faces: characteristic boundary
//                        state, polynomial-correction reconstruction, flux
//                        and lift into the element's face slot
//   time_step_kernel     per-element stable time step and global minimum
//   stage0_kernel        U <- U0 - dt R (buffers swapped by the host)
//
// Reference behaviour followed (VCP3D_CurviLinear/src/):
//   ElementClass::VolumeResidual          ElementClass.cpp:3809-4443
//   ElementClass::{I,J,K}FaceResiduals    ElementClass.cpp:2582-3400
//   FluxesBoundaryFace + NSCBC states     FluxesFace.cpp:16-214,
//                                         BoundaryState.cpp:829-2241
//   AverageDataNSCBC / AverageMachNSCBC   SolverClass.cpp:2502-2581,
//                                         ElementClass.cpp:365-479
//   sponge layer / CML sources            ElementClass.cpp:3887-4117, :4450-4489
//   FluxesInternalFace / InviscidFluxesFace / ViscousFluxesFace /
//   PenAndSymFluxesSymmVar                FluxesFace.cpp:220-343, :813-975,
//                                         :1290-1651, :1152-1284
//   CG_EntropyVarMassMatrix               CG_MassMatrix.cpp:250-504
//   ElementClass::ComputeTimeStep         ElementClass.cpp:1522-1644
//   SolverClass::ExecuteTimeSynchronizationStep  SolverClass.cpp:2177-2422
#include "kernels.hpp"

#include <cstdint>
#include <mutex>

// Fixed device array that vcp_residual_static uses for the face-slot buffer
// instead of the context's own (for callers that cannot allocate device
// memory). Host code never names it directly: its per-device address comes
// from hipGetSymbolAddress (static_scratch).
static __device__ __align__(256) double scratch[vcp::kStaticScratchBytes / sizeof(double)];

namespace vcp {
namespace {

enum : int { kBCOutflowNSCBC = 1, kBCInflowStaticNSCBC = 2, kBCInflowTotalNSCBC = 3, kBCInflowSubsonic = 4,
              kBCOutflowSubsonic = 5, kBCWallHeatFlux = 6, kBCWallIsothermal = 7, kBCWallInviscid = 8,
              kBCInflowSupersonic = 9, kBCOutflowSupersonic = 10, kBCFarfield = 11 };
constexpr double kGam = 1.4;
constexpr double kGm1 = 0.4;
constexpr double kPrLam = 0.72;
constexpr double kPrTurb = 0.90;       // VCP3D_CurviLinear.hpp:176
constexpr double kLamOverMu = -2.0 / 3.0;
constexpr double kInvGm1 = 1.0 / kGm1;

// ------------------------------------------------------------ utilities
constexpr int cmax(int x, int y) { return x > y ? x : y; }
constexpr int cmin(int x, int y) { return x < y ? x : y; }

__device__ inline double wave_sum(double v)
{
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ inline double wave_max(double v)
{
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

// Wave64 reduction through DPP row shifts and row broadcasts (no LDS):
// sum, or max of non-negative values; lanes without a source contribute 0.
// The result (lane 63) is returned to every lane.
template <int CTRL, int ROW_MASK>
__device__ inline double dpp_f64(double v)
{
  const int lo = __double2loint(v), hi = __double2hiint(v);
  const int rlo = __builtin_amdgcn_update_dpp(0, lo, CTRL, ROW_MASK, 0xf, false);
  const int rhi = __builtin_amdgcn_update_dpp(0, hi, CTRL, ROW_MASK, 0xf, false);
  return __hiloint2double(rhi, rlo);
}
// Row shift with every row enabled: lanes shifted in from outside the row
// read 0 (bound_ctrl), so no zeroed destination is needed.
template <int CTRL>
__device__ inline double dpp_shr_f64(double v)
{
  const int lo = __double2loint(v), hi = __double2hiint(v);
  const int rlo = __builtin_amdgcn_mov_dpp(lo, CTRL, 0xf, 0xf, true);
  const int rhi = __builtin_amdgcn_mov_dpp(hi, CTRL, 0xf, 0xf, true);
  return __hiloint2double(rhi, rlo);
}
template <bool MAX>
__device__ inline double wave_reduce_dpp(double v)
{
  auto op = [](double x, double y) { return MAX ? fmax(x, y) : x + y; };
  v = op(v, dpp_shr_f64<0x111>(v));    // row_shr:1
  v = op(v, dpp_shr_f64<0x112>(v));    // row_shr:2
  v = op(v, dpp_shr_f64<0x114>(v));    // row_shr:4
  v = op(v, dpp_shr_f64<0x118>(v));    // row_shr:8   -> lane 15 of each row: row total
  v = op(v, dpp_f64<0x142, 0xa>(v));   // row_bcast:15 into rows 1, 3
  v = op(v, dpp_f64<0x143, 0xc>(v));   // row_bcast:31 into rows 2, 3 -> lane 63: total
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), 63);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), 63);
  return __hiloint2double(hi, lo);
}

// Block reduction with one barrier: partial results go to alternating
// halves of red (2 * 3 * NT/64 doubles); every half starts at the same offset
// whatever NV is, so a half is rewritten only after every thread has passed
// the barrier of the following reduction, also when consecutive reductions
// carry different numbers of values. The max slots (q >= NS) take
// non-negative values. Only waves [W0, W0 + NW) hold contributions; the
// others skip the wave reduction.
template <int NT, int NV, int NS, int NW = NT / 64, int W0 = 0>
__device__ inline void block_reduce1(double (&v)[NV], double *red, int &half)
{
  static_assert(NV >= 1 && NV <= 3, "a half holds at most three values per wave");
  static_assert(W0 >= 0 && W0 + NW <= NT / 64, "contributing waves");
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  double *buf = red + half * 3 * (NT / 64);
  if (wid >= W0 && wid < W0 + NW) {
#pragma unroll
    for (int q = 0; q < NV; ++q) v[q] = (q < NS) ? wave_reduce_dpp<false>(v[q]) : wave_reduce_dpp<true>(v[q]);
    if (lane == 0)
#pragma unroll
      for (int q = 0; q < NV; ++q) buf[wid * NV + q] = v[q];
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    double s = buf[W0 * NV + q];
#pragma unroll
    for (int w = W0 + 1; w < W0 + NW; ++w) s = (q < NS) ? s + buf[w * NV + q] : fmax(s, buf[w * NV + q]);
    v[q] = s;
  }
  half ^= 1;
}

// t^(-1/(gamma-1)) = t^(-5/2) for gamma = 1.4, as 1 / (t^2 sqrt(t)).
__device__ inline double pow_m2p5(double t) { return 1.0 / (t * t * sqrt(t)); }

__device__ inline void ent2prim(const double V[5], double &rho, double &u, double &v, double &w,
                                double &p, double &eKin)
{
  const double V4Inv = 1.0 / V[4];
  u = -V4Inv * V[1]; v = -V4Inv * V[2]; w = -V4Inv * V[3];
  eKin = 0.5 * (u * u + v * v + w * w);
  const double s = kGam - kGm1 * (V[0] - V[4] * eKin);
  rho = pow_m2p5(-V[4] * exp(s));
  p = -rho * V4Inv;
}


// Eddy viscosity mu_SGS = rho nu_t of the subgrid scale model at one point,
// from the velocity gradient G[i][j] = du_i/dx_j and the LES length scale len.
// WALE (CWALEModel::EddyViscosity, SGSModel.cpp:34-113): with S the
// symmetric part of G and Sd the traceless symmetric part of G^2,
// nu_t = (0.325 len)^2 |Sd|^3 / max(|S|^5 + |Sd|^(5/2), 1e-30), |X|^2 = X_ij X_ij.
// Vreman (CVremanModel::EddyViscosity, :128-210): alpha = G with its trace
// taken off the diagonal, beta = len^2 alpha alpha^T, B the sum of the
// principal 2 x 2 minors of beta (clipped at 0),
// nu_t = 0.07 sqrt(B / max(|alpha|^2, 1e-30)).
__device__ inline double eddy_viscosity(int model, double len, double rho, const double G[3][3])
{
  constexpr double third = 1.0 / 3.0, eps = 1e-30;
  if (model == kSgsWale) {
    double G2[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) G2[i][j] = G[i][0] * G[0][j] + G[i][1] * G[1][j] + G[i][2] * G[2][j];
    const double tr = third * (G2[0][0] + G2[1][1] + G2[2][2]);
    double s2 = 0.0, sd2 = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const double d = G2[i][i] - tr;
      s2 += G[i][i] * G[i][i];
      sd2 += d * d;
#pragma unroll
      for (int j = i + 1; j < 3; ++j) {
        const double sij = 0.5 * (G[i][j] + G[j][i]), dij = 0.5 * (G2[i][j] + G2[j][i]);
        s2 += 2.0 * sij * sij;
        sd2 += 2.0 * dij * dij;
      }
    }
    const double sd32 = sd2 * sqrt(sd2);
    const double den = s2 * s2 * sqrt(s2) + sqrt(sd32 * sd2);
    const double l = 0.325 * len;
    return rho * (l * l) * sd32 / fmax(den, eps);
  }
  if (model == kSgsVreman) {
    const double tr = third * (G[0][0] + G[1][1] + G[2][2]);
    double al[3][3], a2 = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        al[i][j] = i == j ? G[i][j] - tr : G[i][j];
        a2 += al[i][j] * al[i][j];
      }
    const double l2 = len * len;
    double be[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = i; j < 3; ++j) be[i][j] = l2 * (al[i][0] * al[j][0] + al[i][1] * al[j][1] + al[i][2] * al[j][2]);
    const double B = be[0][0] * be[1][1] - be[0][1] * be[0][1] + be[0][0] * be[2][2] - be[0][2] * be[0][2]
                   + be[1][1] * be[2][2] - be[1][2] * be[1][2];
    return rho * 0.07 * sqrt(fmax(B, 0.0) / fmax(a2, eps));
  }
  return 0.0;
}

// Eddy viscosity at a face point (FluxesFace.cpp:1409-1459, :1481-1494): the
// model on the density of the averaged entropy variables and the velocity
// gradients of the averaged state and the averaged Cartesian gradients
// gx, gy, gz, with the face's LES length scale. Evaluated before the fluxes,
// so that its temporaries are dead where those peak in registers.
__device__ __forceinline__ double face_eddy(int sgs, double lenLES, const double VL[5], const double VR[5],
                                            const double gx[5], const double gy[5], const double gz[5])
{
  const double V0a = 0.5 * (VL[0] + VR[0]), V4a = 0.5 * (VL[4] + VR[4]);
  const double V4Inv = 2.0 / (VL[4] + VR[4]);
  const double u = -V4Inv * 0.5 * (VL[1] + VR[1]), v = -V4Inv * 0.5 * (VL[2] + VR[2]),
               w = -V4Inv * 0.5 * (VL[3] + VR[3]);
  const double ek = 0.5 * (u * u + v * v + w * w);
  const double rhoA = pow_m2p5(-V4a * exp(kGam - kGm1 * (V0a - V4a * ek)));
  const double vel[3] = {u, v, w};
  double G[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    G[i][0] = -V4Inv * (gx[1 + i] + vel[i] * gx[4]);
    G[i][1] = -V4Inv * (gy[1 + i] + vel[i] * gy[4]);
    G[i][2] = -V4Inv * (gz[1 + i] + vel[i] * gz[4]);
  }
  return eddy_viscosity(sgs, lenLES, rhoA, G);
}

// Ismail-Roe + viscous + interior-penalty + symmetrising fluxes at one face
// point: VL/VR entropy states (L -> R along nn), gx/gy/gz the averaged
// Cartesian gradients, wA = weight * area factor. fl: normal flux, FX/FY/FZ:
// Cartesian symmetrising fluxes (FluxesFace.cpp:813-975, :1290-1651,
// :1152-1284). BND: the boundary form, where factU (wallPerm, 0 on walls)
// multiplies the convective and dissipation parts of the inviscid flux and a
// prescribed wall heat flux (hf) zeroes the laminar heat-flux factor and adds
// wA * qWall to the energy flux (FluxesFace.cpp:1312-1318, :1586-1593).
// sgs: with a subgrid scale model, its eddy viscosity ev at the point
// (face_eddy) adds to mu in the stresses, the penalty and the symmetrising
// fluxes, with the turbulent Prandtl number in the heat flux
// (FluxesFace.cpp:1537-1554, :1096-1103, :1198-1213).
template <bool BND = false>
__device__ __forceinline__ void point_flux(const double VL[5], const double VR[5], const double gx[5],
                                           const double gy[5], const double gz[5], const double nn[3],
                                           double wA, double tau, double mu, double thetaHalf,
                                           double fl[5], double FX[5], double FY[5], double FZ[5],
                                           double factU = 1.0, bool hf = false, double qWall = 0.0,
                                           double *aux = nullptr, bool sgs = false, double ev = 0.0)
{
  const double factHF = (BND && hf) ? 0.0 : kGam / kPrLam;
  const double factHFT = (BND && hf) ? 0.0 : kGam / kPrTurb;
  // total viscosity, total heat-flux factor and their ratio
  double muT = mu, fhT = mu * factHF, fhR = factHF;
  const double lam = kLamOverMu, lamp1 = lam + 1.0, lamp2 = lam + 2.0;
  {
    double rL, uL, vL, wL, pL, eL, rR, uR, vR, wR, pR, eR;
    ent2prim(VL, rL, uL, vL, wL, pL, eL);
    ent2prim(VR, rR, uR, vR, wR, pR, eR);
    if (BND && aux) aux[0] = pL + pR;   // for the wall forces
    const double dV0 = VR[0] - VL[0], dV1 = VR[1] - VL[1], dV2 = VR[2] - VL[2],
                 dV3 = VR[3] - VL[3], dV4 = VR[4] - VL[4];
    // sqrt(rho / p) = sqrt(-V4)
    const double z0L = sqrt(-VL[4]), z0R = sqrt(-VR[4]);
    const double z4L = sqrt(rL * pL), z4R = sqrt(rR * pR);
    const double z0A = 0.5 * (z0L + z0R), z4A = 0.5 * (z4L + z4R);
    const double z1A = 0.5 * (uL * z0L + uR * z0R), z2A = 0.5 * (vL * z0L + vR * z0R),
                 z3A = 0.5 * (wL * z0L + wR * z0R);
    // logarithmic means: F = log(zeta) / (2 f), f = (zeta - 1) / (zeta + 1)
    // = (zL - zR) / (zL + zR), by its series for small f
    constexpr double c3 = 1.0 / 3.0, c5 = 1.0 / 5.0, c7 = 1.0 / 7.0, c9 = 1.0 / 9.0;
    double f = (z0L - z0R) / (z0L + z0R), uu = f * f, F;
    F = (uu < 0.01) ? 1.0 + uu * (c3 + uu * (c5 + uu * (c7 + uu * c9))) : log(z0L / z0R) / (2.0 * f);
    const double z0Log = z0A / F;
    f = (z4L - z4R) / (z4L + z4R); uu = f * f;
    F = (uu < 0.01) ? 1.0 + uu * (c3 + uu * (c5 + uu * (c7 + uu * c9))) : log(z4L / z4R) / (2.0 * f);
    const double z4Log = z4A / F;
    const double oz0 = 1.0 / z0A;
    const double rA = z0A * z4Log, p1 = oz0 * z4A, orA = 1.0 / rA;
    const double p2 = 0.5 * ((kGam + 1.0) / kGam * z4Log / z0Log + kGm1 / kGam * p1);
    const double uA = oz0 * z1A, vA = oz0 * z2A, wAv = oz0 * z3A;
    const double vn = uA * nn[0] + vA * nn[1] + wAv * nn[2];
    const double kin = 0.5 * (uA * uA + vA * vA + wAv * wAv);
    const double a2 = kGam * p2 * orA, aa = sqrt(a2), oa = 1.0 / aa;
    const double H = a2 * kInvGm1 + kin, E = H - p2 * orA;
    double abv1 = rA * (uA * dV1 + vA * dV2 + wAv * dV3);
    double abv2 = abv1 + rA * (dV0 + H * dV4);
    const double dr = abv1 + rA * (dV0 + E * dV4);
    const double dru = uA * abv2 + p2 * dV1, drv = vA * abv2 + p2 * dV2, drw = wAv * abv2 + p2 * dV3;
    const double drE = H * abv1 + rA * E * (dV0 + H * dV4) + p2 * kin * dV4;
    const double fac = 0.5 * (1.0 + 1.0 / 6.0);
    const double lam1 = fabs(vn + aa) * fac, lam2 = fabs(vn - aa) * fac, lam3 = 0.5 * fabs(vn);
    abv1 = 0.5 * (lam1 + lam2); abv2 = 0.5 * (lam1 - lam2);
    const double abv3 = abv1 - lam3;
    const double abv4 = kGm1 * (kin * dr - uA * dru - vA * drv - wAv * drw + drE);
    const double abv5 = nn[0] * dru + nn[1] * drv + nn[2] * drw - vn * dr;
    const double abv6 = abv3 * abv4 * (oa * oa) + abv2 * abv5 * oa;
    const double abv7 = abv2 * abv4 * oa + abv3 * abv5;
    if constexpr (BND) {
      fl[0] = wA * (factU * rA * vn - factU * (lam3 * dr + abv6));
      fl[1] = wA * (factU * rA * vn * uA + p1 * nn[0] - factU * (lam3 * dru + uA * abv6 + nn[0] * abv7));
      fl[2] = wA * (factU * rA * vn * vA + p1 * nn[1] - factU * (lam3 * drv + vA * abv6 + nn[1] * abv7));
      fl[3] = wA * (factU * rA * vn * wAv + p1 * nn[2] - factU * (lam3 * drw + wAv * abv6 + nn[2] * abv7));
      fl[4] = wA * (factU * rA * vn * H - factU * (lam3 * drE + H * abv6 + vn * abv7));
    } else {
      fl[0] = wA * (rA * vn - (lam3 * dr + abv6));
      fl[1] = wA * (rA * vn * uA + p1 * nn[0] - (lam3 * dru + uA * abv6 + nn[0] * abv7));
      fl[2] = wA * (rA * vn * vA + p1 * nn[1] - (lam3 * drv + vA * abv6 + nn[1] * abv7));
      fl[3] = wA * (rA * vn * wAv + p1 * nn[2] - (lam3 * drw + wAv * abv6 + nn[2] * abv7));
      fl[4] = wA * (rA * vn * H - (lam3 * drE + H * abv6 + vn * abv7));
    }
  }
  const double V4Inv = 2.0 / (VL[4] + VR[4]);
  const double u = -V4Inv * 0.5 * (VL[1] + VR[1]), vv = -V4Inv * 0.5 * (VL[2] + VR[2]),
               w = -V4Inv * 0.5 * (VL[3] + VR[3]);
  {
    const double dudx = -V4Inv * (gx[1] + u * gx[4]), dvdx = -V4Inv * (gx[2] + vv * gx[4]), dwdx = -V4Inv * (gx[3] + w * gx[4]);
    const double dudy = -V4Inv * (gy[1] + u * gy[4]), dvdy = -V4Inv * (gy[2] + vv * gy[4]), dwdy = -V4Inv * (gy[3] + w * gy[4]);
    const double dudz = -V4Inv * (gz[1] + u * gz[4]), dvdz = -V4Inv * (gz[2] + vv * gz[4]), dwdz = -V4Inv * (gz[3] + w * gz[4]);
    const double fe = V4Inv * V4Inv * kInvGm1;
    const double dedx = fe * gx[4], dedy = fe * gy[4], dedz = fe * gz[4];
    if (sgs) {
      muT = mu + ev;
      fhT = mu * factHF + ev * factHFT;
      fhR = fhT / muT;
    }
    const double nx = wA * nn[0], ny = wA * nn[1], nz = wA * nn[2];
    const double lamDiv = kLamOverMu * (dudx + dvdy + dwdz);
    const double txx = muT * (2.0 * dudx + lamDiv), tyy = muT * (2.0 * dvdy + lamDiv), tzz = muT * (2.0 * dwdz + lamDiv);
    const double txy = muT * (dudy + dvdx), txz = muT * (dudz + dwdx), tyz = muT * (dvdz + dwdy);
    const double qx = -fhT * dedx, qy = -fhT * dedy, qz = -fhT * dedz;
    fl[1] -= txx * nx + txy * ny + txz * nz;
    fl[2] -= txy * nx + tyy * ny + tyz * nz;
    fl[3] -= txz * nx + tyz * ny + tzz * nz;
    fl[4] -= (u * txx + vv * txy + w * txz - qx) * nx + (u * txy + vv * tyy + w * tyz - qy) * ny
           + (u * txz + vv * tyz + w * tzz - qz) * nz;
    if (BND && hf) fl[4] += wA * qWall;
    if (BND && aux) { aux[1] = txx; aux[2] = tyy; aux[3] = tzz; aux[4] = txy; aux[5] = txz; aux[6] = tyz; }
  }
  const double pOvRho = -V4Inv, eInt = pOvRho * kInvGm1;
  const double dV1 = VL[1] - VR[1], dV2 = VL[2] - VR[2], dV3 = VL[3] - VR[3], dV4 = VL[4] - VR[4];
  const double factFlux = pOvRho * wA * muT;
  const double abvE = u * u + vv * vv + w * w + fhR * eInt;
  const double un = u * nn[0] + vv * nn[1] + w * nn[2];
  const double dVn = dV1 * nn[0] + dV2 * nn[1] + dV3 * nn[2];
  const double ttau = tau * factFlux, abvMom = lamp1 * (dVn + un * dV4);
  fl[1] += ttau * (abvMom * nn[0] + dV1 + u * dV4);
  fl[2] += ttau * (abvMom * nn[1] + dV2 + vv * dV4);
  fl[3] += ttau * (abvMom * nn[2] + dV3 + w * dV4);
  fl[4] += ttau * (lamp1 * un * dVn + u * dV1 + vv * dV2 + w * dV3 + (abvE + lamp1 * un * un) * dV4);
  const double sx = thetaHalf * factFlux * nn[0], sy = thetaHalf * factFlux * nn[1], sz = thetaHalf * factFlux * nn[2];
  const double K11_14 = u * lamp2, K11_44 = abvE + u * u * lamp1;
  const double K22_24 = vv * lamp2, K22_44 = abvE + vv * vv * lamp1;
  const double K33_34 = w * lamp2, K33_44 = abvE + w * w * lamp1;
  const double K12_14 = vv * lam, K12_42 = u * lam, K12_44 = u * vv * lamp1;
  const double K13_14 = w * lam, K13_43 = u * lam, K13_44 = u * w * lamp1;
  const double K23_24 = w * lam, K23_43 = vv * lam, K23_44 = vv * w * lamp1;
  FX[0] = 0.0;
  FX[1] = -(lamp2 * dV1 + K11_14 * dV4) * sx - (lam * dV2 + K12_14 * dV4) * sy - (lam * dV3 + K13_14 * dV4) * sz;
  FX[2] = -(dV2 + vv * dV4) * sx - (dV1 + u * dV4) * sy;
  FX[3] = -(dV3 + w * dV4) * sx - (dV1 + u * dV4) * sz;
  FX[4] = -(K11_14 * dV1 + vv * dV2 + w * dV3 + K11_44 * dV4) * sx
          - (vv * dV1 + K12_42 * dV2 + K12_44 * dV4) * sy - (w * dV1 + K13_43 * dV3 + K13_44 * dV4) * sz;
  FY[0] = 0.0;
  FY[1] = -(dV2 + vv * dV4) * sx - (dV1 + u * dV4) * sy;
  FY[2] = -(lam * dV1 + K12_42 * dV4) * sx - (lamp2 * dV2 + K22_24 * dV4) * sy - (lam * dV3 + K23_24 * dV4) * sz;
  FY[3] = -(dV3 + w * dV4) * sy - (dV2 + vv * dV4) * sz;
  FY[4] = -(K12_14 * dV1 + u * dV2 + K12_44 * dV4) * sx
          - (u * dV1 + K22_24 * dV2 + w * dV3 + K22_44 * dV4) * sy - (w * dV2 + K23_43 * dV3 + K23_44 * dV4) * sz;
  FZ[0] = 0.0;
  FZ[1] = -(dV3 + w * dV4) * sx - (dV1 + u * dV4) * sz;
  FZ[2] = -(dV3 + w * dV4) * sy - (dV2 + vv * dV4) * sz;
  FZ[3] = -(lam * dV1 + K13_43 * dV4) * sx - (lam * dV2 + K23_43 * dV4) * sy - (lamp2 * dV3 + K33_34 * dV4) * sz;
  FZ[4] = -(K13_14 * dV1 + u * dV3 + K13_44 * dV4) * sx
          - (K23_24 * dV2 + vv * dV3 + K23_44 * dV4) * sy - (u * dV1 + vv * dV2 + K33_34 * dV3 + K33_44 * dV4) * sz;
}

template <int K, int M, typename T = double>
struct FaceOps {
  T A[M * K], D[M * K];        // phi_q(r_p), phi'_q(r_p), row-major (p, q)
  T F[4][K];                   // phi(-1), phi(+1), phi'(-1), phi'(+1)
  double w[M];                 // 1D Gauss weights
  double c03;                  // phi_0^3, the value of the constant mode
};

// Face kernel arguments: the case part (single-case launch) or the case
// table and work items (batched launch), and the operators (the
// metric-indexing mode is the kernel's BUG parameter).
template <int K, int M, typename T = double>
struct FaceArgs {
  CaseDev c;
  const CaseDev *cases;
  const WorkItem *work;
  int w0;
  FaceOps<K, M, T> op;
};

// ============================================================ face kernel
// One wave64 per element min face in direction DIR. Side 0 is the element
// itself (its min face), side 1 the neighbour L (its max face); the normal
// points from L to the element. When the fast contraction (2 * 5 K columns)
// and the points (M^2) each fit the 64 lanes in a single pass, those phases
// overwrite their own input in place (all lanes load before any lane
// stores); otherwise (K = 7, M = 9) they loop over their items and T gets a
// region of its own. LDS regions:
//   X: normal contraction B -> fast contraction T -> flux and Cartesian
//      symmetrising fluxes at the points (over B, or over T when in place)
//   Y: the first lift stage of one side, in two halves
template <int K, int M, typename T = double>
struct FaceSmem {
  static constexpr bool kInPlace = 2 * 5 * K <= 64 && M * M <= 64;
  static constexpr int kB = 2 * 2 * 5 * K * K, kT = 2 * 3 * 5 * K * M, kP = 20 * M * M;
  static constexpr int kOffT = kInPlace ? 0 : cmax(kB, kP);   // the fluxes start at 0
  static constexpr int kX = kInPlace ? cmax(cmax(kB, kT), kP) : kOffT + kT;
  static constexpr int kY = 2 * 5 * M * K;
  static constexpr int kBytes = (int) sizeof(T) * (2 * M * K + 4 * K + kX + kY) + 8 * (M * M + 90 + 10);
  // waves per SIMD to compile for: three when 12 faces fit a CU's 160 KB
#ifndef VCP_F32_FACE_WAVES
#define VCP_F32_FACE_WAVES 3
#endif
  static constexpr int kWaves = sizeof(T) == 4 && 4 * VCP_F32_FACE_WAVES * kBytes <= 160 * 1024 ? VCP_F32_FACE_WAVES
                                : 12 * kBytes <= 160 * 1024 ? 3 : 2;
  T A[M * K], D[M * K];
  double W2[M * M];
  T F[4][K];                       // phi(-1), phi(+1), phi'(-1), phi'(+1)
  double Cm[2][5][9];
  double Mean[2][5];               // float form: mode 0 of both elements
  T X[kX];
  T Y[kY];
};

// args: the kernel's argument block, read in place from the kernel-argument
// segment by both forms, so that run-time-indexed operator rows come through
// scalar loads and the single-case and batched kernels do the same
// arithmetic. T: the type of the contractions (double, or float for the
// mixed-precision form, where the traces carry the variation about each
// element's mean, added back at the points in double, and the pointwise
// fluxes stay in double). SGS: the case has a subgrid scale model (a
// separate instantiation: its eddy-viscosity code raises the register peak).
template <int K, int M, int DIR, bool BUG, bool BATCH, typename T, bool SGS>
__device__ __forceinline__ void face_body(const FaceArgs<K, M, T> &args, const CaseDev a, FaceSmem<K, M, T> &sm,
                                          int e)
{
#define VCP_FOPS (args.op)
  constexpr bool kMixed = sizeof(T) == 4;
  const int L = a.nbr[e * 6 + 2 * DIR];
  if (L < 0) return;  // physical boundary face
  constexpr int ND = K * K * K, N2 = M * M;
  using S = FaceSmem<K, M, T>;
  static_assert(5 * M <= 64 && 5 * K <= 64, "single-pass lift stages");
  // strides of the (normal, fast, slow) axes in the [k][j][i] DOF layout
  constexpr int SN = DIR == 0 ? 1 : DIR == 1 ? K : K * K;
  constexpr int SFA = DIR == 0 ? K : 1;
  constexpr int SSL = DIR == 2 ? K : K * K;
  // parametric slot of each axis: r = 0, s = 1, t = 2 -> (normal, fast, slow)
  constexpr int PN = DIR, PF = DIR == 0 ? 1 : 0, PS = DIR == 2 ? 1 : 2;
  T *sA = sm.A, *sD = sm.D;
  double *sW2 = sm.W2;
  T *B = sm.X;                                // [side][value|deriv][5][K][K]
  T *Tb = sm.X + S::kOffT;                    // [side][A|D|N][5][K][M]
  T *sFlux = sm.X, *sFC = sm.X + 5 * N2;      // flux [5][N2], Cartesian sym [x|y|z][5][N2]
  const int tid = threadIdx.x;
  // operators in LDS: read as wave-uniform LDS operands in the contractions
  // (as kernel arguments they would not fit the scalar registers)
  for (int i = tid; i < M * K; i += 64) { sA[i] = VCP_FOPS.A[i]; sD[i] = VCP_FOPS.D[i]; }
  for (int i = tid; i < 4 * K; i += 64) sm.F[i / K][i % K] = VCP_FOPS.F[i / K][i % K];
  const T *A = sA, *D = sD;
  // operators of the pencil contractions: the kernel arguments (wave-uniform
  // scalar loads) while the argument block is under 1 KB; a larger one
  // (K >= 6) the compiler copies to scratch when it is indexed at run time,
  // so those read the LDS copies instead
  constexpr bool kArgOps = sizeof(FaceOps<K, M, T>) < 1024;
  const T *KA = kArgOps ? VCP_FOPS.A : sA, *KD = kArgOps ? VCP_FOPS.D : sD;
  for (int i = tid; i < N2; i += 64) sW2[i] = VCP_FOPS.w[i % M] * VCP_FOPS.w[i / M];
  const double *g = a.fgeo + ((size_t) e * 3 + DIR) * kFaceGeo * N2;
  // Loads issued here and used after the contractions, so that their latency
  // passes under them: the two length scales and, when the points fit the
  // wave in one pass, this lane's point metric (normal, area and, unless the
  // metric comes from sm.Cm, both sides' inverse metric).
  const double lenE = a.econst[(size_t) e * kElemConst + DIR], lenL = a.econst[(size_t) L * kElemConst + DIR];
  constexpr bool kGeoEarly = N2 <= 64;
  double geoCm[2][9], geoN[4];
  if constexpr (kGeoEarly) {
    const int mm = tid < N2 ? tid : 0;
    if constexpr (!BUG)
#pragma unroll
      for (int q = 0; q < 9; ++q) { geoCm[0][q] = g[(13 + q) * N2 + mm]; geoCm[1][q] = g[(4 + q) * N2 + mm]; }
#pragma unroll
    for (int q = 0; q < 4; ++q) geoN[q] = g[q * N2 + mm];
  }
  if constexpr (BUG)
#pragma unroll
    for (int it = 0; it < 2; ++it) {
      const int i = tid + 64 * it;
      if (i < 2 * 5 * 9) {
        const int side = i / 45, l = (i / 9) % 5, q = i % 9;
        const int pt = l < N2 ? l : 0;
        sm.Cm[side][l][q] = g[(side ? 4 + q : 13 + q) * N2 + pt];
      }
    }
  if constexpr (kMixed)
    if (tid < 10) sm.Mean[tid / 5][tid % 5] = a.U[(size_t) (tid < 5 ? e : L) * 5 * ND + (tid % 5) * ND];
  __syncthreads();
  // normal contraction straight from HBM: B[side][0|1][v][s][f]; the loads of
  // both sides (K <= 6, at most 40 registers) or of one side at a time (K = 7)
  // are issued before the first of them is waited for. In the float form the
  // mean (mode 0) is added at the points.
  constexpr int NIT = (5 * K * K + 63) / 64;
  constexpr int SB = 2 * NIT * K <= 40 ? 2 : 1;
#pragma unroll
  for (int s0 = 0; s0 < 2; s0 += SB) {
    double xu[SB][NIT][K];
#pragma unroll
    for (int sb = 0; sb < SB; ++sb) {
      const double *Us = a.U + (size_t) (s0 + sb ? L : e) * 5 * ND;
#pragma unroll
      for (int it = 0; it < NIT; ++it) {
        const int pc0 = tid + 64 * it, pc = pc0 < 5 * K * K ? pc0 : 0;
        const int v = pc / (K * K), sl = (pc / K) % K, f = pc % K;
        const double *u = Us + v * ND + sl * SSL + f * SFA;
#pragma unroll
        for (int t = 0; t < K; ++t) xu[sb][it][t] = u[t * SN];
      }
    }
#pragma unroll
    for (int sb = 0; sb < SB; ++sb) {
      const int side = s0 + sb;
      const T *Fv = sm.F[side], *dFv = sm.F[2 + side];
#pragma unroll
      for (int it = 0; it < NIT; ++it) {
        const int pc = tid + 64 * it;
        if (pc < 5 * K * K) {
          const int sl = (pc / K) % K, f = pc % K;
          T b0 = 0, b1 = 0;
#pragma unroll
          for (int t = 0; t < K; ++t) {
            const T x = (kMixed && t == 0 && sl == 0 && f == 0) ? T(0) : T(xu[sb][it][t]);
            b0 += Fv[t] * x; b1 += dFv[t] * x;
          }
          B[(side * 2 + 0) * 5 * K * K + pc] = b0;
          B[(side * 2 + 1) * 5 * K * K + pc] = b1;
        }
      }
    }
  }
  __syncthreads();
  // fast tangential contraction (in place over B when single-pass):
  // T[side][A|D|N][v][s][f'], f' < M
  for (int c0 = 0; c0 < 2 * 5 * K; c0 += 64) {
    const int c = c0 + tid;
    if (c >= 2 * 5 * K) break;
    const int side = c / (5 * K), vs = c - side * 5 * K;
    T x0[K], x1[K];
#pragma unroll
    for (int f = 0; f < K; ++f) {
      x0[f] = B[(side * 2 + 0) * 5 * K * K + vs * K + f];
      x1[f] = B[(side * 2 + 1) * 5 * K * K + vs * K + f];
    }
    T *Ts = Tb + side * 3 * 5 * K * M;
#pragma unroll 1
    for (int n = 0; n < M; ++n) {
      T tA = 0, tD = 0, tN = 0;
#pragma unroll
      for (int f = 0; f < K; ++f) { tA += KA[n * K + f] * x0[f]; tD += KD[n * K + f] * x0[f]; tN += KA[n * K + f] * x1[f]; }
      Ts[vs * M + n] = tA; Ts[5 * K * M + vs * M + n] = tD; Ts[2 * 5 * K * M + vs * M + n] = tN;
    }
  }
  __syncthreads();
  // per point: traces, averaged Cartesian gradients of the entropy variables
  // and the fluxes (Ismail-Roe, viscous, penalty, symmetrising); the outputs
  // overwrite T in place when single-pass, else the dead B
  const double lenScale = fmin(lenE, lenL);
  const double tau = a.tauFact / lenScale;
  // LES length scale of the face: the mean of the two elements'
  // (ElementClass.cpp:2655)
  const double lenLES = SGS ? 0.5 * (a.econst[(size_t) e * kElemConst + kElemLenLES]
                                     + a.econst[(size_t) L * kElemConst + kElemLenLES]) : 0.0;
  double evMax = 0.0;
  for (int m0 = 0; m0 < N2; m0 += 64) {
    const int m = m0 + tid;
    if (m >= N2) break;
    const int sp = m / M, fp = m - sp * M;
    T ai[K], di[K];
#pragma unroll
    for (int sl = 0; sl < K; ++sl) { ai[sl] = sA[sp * K + sl]; di[sl] = sD[sp * K + sl]; }
    double Vs[2][5], gx[5], gy[5], gz[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) { gx[v] = 0.0; gy[v] = 0.0; gz[v] = 0.0; }
#pragma unroll
    for (int side = 0; side < 2; ++side) {
      double cm[9];
      if constexpr (!BUG) {
#pragma unroll
        for (int q = 0; q < 9; ++q) cm[q] = kGeoEarly ? geoCm[side][q] : g[(side ? 4 + q : 13 + q) * N2 + m];
      }
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        const T *tA = Tb + (side * 3 + 0) * 5 * K * M + v * K * M + fp;
        const T *tD = Tb + (side * 3 + 1) * 5 * K * M + v * K * M + fp;
        const T *tN = Tb + (side * 3 + 2) * 5 * K * M + v * K * M + fp;
        T val = 0, gN = 0, gF = 0, gS = 0;
#pragma unroll
        for (int sl = 0; sl < K; ++sl) {
          const T xa = tA[sl * M];
          val += ai[sl] * xa; gS += di[sl] * xa; gN += ai[sl] * tN[sl * M]; gF += ai[sl] * tD[sl * M];
        }
        Vs[side][v] = kMixed ? double(val) + VCP_FOPS.c03 * sm.Mean[side][v] : double(val);
        const double gr = DIR == 0 ? gN : gF;
        const double gs = DIR == 0 ? gF : DIR == 1 ? gN : gS;
        const double gt = DIR == 2 ? gN : gS;
        const double *c = cm;
        if constexpr (BUG) c = sm.Cm[side][v];
        gx[v] += 0.5 * (gr * c[0] + gs * c[3] + gt * c[6]);
        gy[v] += 0.5 * (gr * c[1] + gs * c[4] + gt * c[7]);
        gz[v] += 0.5 * (gr * c[2] + gs * c[5] + gt * c[8]);
      }
    }
    // materialise the traces and gradients here: otherwise the compiler sinks
    // their computation into the flux and keeps the LDS operands live
#pragma unroll
    for (int v = 0; v < 5; ++v) asm volatile("" : "+v"(Vs[0][v]), "+v"(Vs[1][v]), "+v"(gx[v]), "+v"(gy[v]), "+v"(gz[v]));
    const double nn[3] = {kGeoEarly ? geoN[0] : g[0 * N2 + m], kGeoEarly ? geoN[1] : g[1 * N2 + m],
                          kGeoEarly ? geoN[2] : g[2 * N2 + m]};
    const double wA = sW2[m] * (kGeoEarly ? geoN[3] : g[3 * N2 + m]);
    double fl[5], FX[5], FY[5], FZ[5], ev = 0.0;
    if constexpr (SGS) {
      ev = face_eddy(a.sgs, lenLES, Vs[1], Vs[0], gx, gy, gz);
      evMax = fmax(evMax, ev);
    }
    point_flux(Vs[1], Vs[0], gx, gy, gz, nn, wA, tau, a.mu, a.thetaHalf, fl, FX, FY, FZ, 1.0, false, 0.0, nullptr,
               SGS, ev);
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      sFlux[v * N2 + m] = T(fl[v]);
      sFC[v * N2 + m] = T(FX[v]); sFC[(5 + v) * N2 + m] = T(FY[v]); sFC[(10 + v) * N2 + m] = T(FZ[v]);
    }
  }
  if (SGS && a.eddyPart) {
    // the face's eddy-viscosity maximum for the monitoring data
    // (FluxesFace.cpp:336-342)
    evMax = wave_max(evMax);
    if (tid == 0) a.eddyPart[a.nElem + 3 * e + DIR] = evMax;
  }
  __syncthreads();
  // ---- lift, one side at a time: fast-axis contraction into Y (first the
  // flux + fast-slot and the slow-slot parts, then the normal-slot part),
  // each followed by the slow-axis contraction straight into the face slot
  // of that side: the value and normal-derivative coefficient planes
  // [0|1][v][slow][fast]; the volume kernel applies the face vectors
#pragma unroll
  for (int side = 0; side < 2; ++side) {
    const T sign = side ? T(1) : T(-1);
    double *out = a.faceRes + ((size_t) (side ? L : e) * 6 + 2 * DIR + side) * 10 * K * K;
    T gv[K];   // value plane of this lane's (v, f) column, over both parts
#pragma unroll
    for (int part = 0; part < 2; ++part) {
      // Cartesian symmetrising flux mapped to this side's parametric slot(s)
      if (tid < 5 * M) {
        const int v = tid / M, sp = tid - v * M;
        T x0[M], x1[M];
#pragma unroll
        for (int q = 0; q < M; ++q) {
          const int m = sp * M + q;
          const double fx = sFC[v * N2 + m], fy = sFC[(5 + v) * N2 + m], fz = sFC[(10 + v) * N2 + m];
          double c[9];
          if constexpr (BUG) {
#pragma unroll
            for (int t = 0; t < 9; ++t) c[t] = sm.Cm[side][v][t];
          } else {
#pragma unroll
            for (int t = 0; t < 9; ++t) c[t] = g[(side ? 4 + t : 13 + t) * N2 + m];
          }
          if (part == 0) {
            x0[q] = sign * sFlux[v * N2 + m];                                      // lifted with A
            x1[q] = T(fx * c[3 * PF] + fy * c[3 * PF + 1] + fz * c[3 * PF + 2]);   // lifted with D
          } else {
            x0[q] = T(fx * c[3 * PS] + fy * c[3 * PS + 1] + fz * c[3 * PS + 2]);   // slow slot
            x1[q] = T(fx * c[3 * PN] + fy * c[3 * PN + 1] + fz * c[3 * PN + 2]);   // normal slot
          }
        }
        T *Ys = sm.Y;
#pragma unroll 1
        for (int n = 0; n < K; ++n) {
          T p0 = 0, p1 = 0;
#pragma unroll
          for (int q = 0; q < M; ++q) {
            if (part == 0) { p0 += KA[q * K + n] * x0[q]; p0 += KD[q * K + n] * x1[q]; }
            else { p0 += KA[q * K + n] * x0[q]; p1 += KA[q * K + n] * x1[q]; }
          }
          Ys[tid * K + n] = p0;
          if (part == 1) Ys[5 * M * K + tid * K + n] = p1;
        }
      }
      __syncthreads();
      // slow axis: part 0 gives the value plane with A (flux + fast slot)
      // and, after part 1, D (slow slot); the normal slot gives the
      // derivative plane with A
      if (tid < 5 * K) {
        const int v = tid / K, f = tid - v * K;
        T x0[M], x1[M];
#pragma unroll
        for (int q = 0; q < M; ++q) {
          const int j = (v * M + q) * K + f;
          x0[q] = sm.Y[j];
          if (part == 1) x1[q] = sm.Y[5 * M * K + j];
        }
#pragma unroll
        for (int n = 0; n < K; ++n) {
          T g0 = 0, g1 = 0;
#pragma unroll
          for (int q = 0; q < M; ++q) {
            if (part == 0) g0 += KA[q * K + n] * x0[q];
            else { g0 += KD[q * K + n] * x0[q]; g1 += KA[q * K + n] * x1[q]; }
          }
          const int o = (v * K + n) * K + f;
          if (part == 0) gv[n] = g0;
          else { out[o] = double(gv[n] + g0); out[5 * K * K + o] = double(g1); }
          asm volatile("" ::: "memory");   // one operator column's scalar loads at a time
        }
      }
      __syncthreads();
    }
  }
#undef VCP_FOPS
}

// The three directions in one launch: blockIdx.y is the direction;
// blockIdx.x the element, or (BATCH) a work item (case, element) of the
// group. The argument block is read through the kernel-argument segment
// pointer: a by-value copy indexed at run time would go to scratch. The SGS
// form is compiled for two waves per SIMD, where its eddy-viscosity code fits
// the registers without spilling.
template <int K, int M, bool BUG, bool BATCH, typename T, bool SGS>
__global__ __launch_bounds__(64, (SGS ? cmin(2, FaceSmem<K, M, T>::kWaves) : FaceSmem<K, M, T>::kWaves))
void face_kernel(FaceArgs<K, M, T> a_)
{
  __shared__ FaceSmem<K, M, T> sm;
  (void) a_;
  const FaceArgs<K, M, T> &a = *(const FaceArgs<K, M, T> *) __builtin_amdgcn_kernarg_segment_ptr();
  CaseDev c;
  int e;
  if constexpr (BATCH) {
    const WorkItem w = a.work[a.w0 + blockIdx.x];
    c = a.cases[w.c];
    e = w.e;
    if (c.flags[kWordActive] == 0) return;
  } else {
    c = a.c;
    e = blockIdx.x;
  }
  if (blockIdx.y == 0) face_body<K, M, 0, BUG, BATCH, T, SGS>(a, c, sm, e);
  else if (blockIdx.y == 1) face_body<K, M, 1, BUG, BATCH, T, SGS>(a, c, sm, e);
  else face_body<K, M, 2, BUG, BATCH, T, SGS>(a, c, sm, e);
}

// 1D operators of the standard element as kernel arguments: indexed with
// compile-time constants after unrolling, they are read into scalar registers.
template <int K, int M, typename T = double>
struct Ops {
  T A[M * K];   // phi_q(r_p), row-major (p, q)
  T D[M * K];   // phi'_q(r_p)
};

// Volume kernel arguments: the case part (single-case launch) or the case
// table and work items (batched launch), the stage control and what depends
// on (K, M) only.
template <int K, int M, typename T = double>
struct VolArgs {
  CaseDev c;
  const CaseDev *cases;
  const WorkItem *work;
  int w0;
  StageCtl ctl;
  TableOffsets to;
  int bug;
  Ops<K, M, T> op;
  double c0[K];                           // phi_q(0): basis at the element centre
  double fvec[4][K];                      // phi(-1), phi(+1), phi'(-1), phi'(+1)
};

// Characteristic-matching-layer source at one point (ElementClass.cpp:
// 3887-4117): wave amplitudes of the Cartesian primitive gradients per
// direction with the outgoing eigenvalues of the layer orientation removed,
// P = S L mapped to the conservative variables. V: entropy variables,
// gx/gy/gz their Cartesian gradients, wgt[d] = -J w sigma_d scale_d,
// nsg[d] the layer orientation. Accumulates into out.
__device__ __forceinline__ void cml_source(const double V[5], const double gx[5], const double gy[5],
                                           const double gz[5], const double wgt[3], const double nsg[3],
                                           double out[5])
{
  const double V4Inv = 1.0 / V[4];
  const double u = -V4Inv * V[1], v = -V4Inv * V[2], w = -V4Inv * V[3];
  const double ek = 0.5 * (u * u + v * v + w * w);
  const double s = kGam - kGm1 * (V[0] - V[4] * ek);
  const double rho = pow_m2p5(-V[4] * exp(s));
  const double p = -rho * V4Inv, rE = p / kGm1 + rho * ek;
  const double H = rE / rho - V4Inv, a2 = -kGam * V4Inv, a = sqrt(a2);
  const double rhoa = rho * a, ovra = 1.0 / rhoa, ova2 = 1.0 / a2, rova = rho / a;
  const double rhou = rho * u, rhov = rho * v, rhow = rho * w;
  const double vel[3] = {u, v, w};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double *G = d == 0 ? gx : d == 1 ? gy : gz;
    double dq[5];
    dq[0] = rho * G[0] + rhou * G[1] + rhov * G[2] + rhow * G[3] + rE * G[4];
    dq[1] = -V4Inv * G[1] + (-V4Inv * u) * G[4];
    dq[2] = -V4Inv * G[2] + (-V4Inv * v) * G[4];
    dq[3] = -V4Inv * G[3] + (-V4Inv * w) * G[4];
    dq[4] = p * G[0] + (p * u) * G[1] + (p * v) * G[2] + (p * w) * G[3] + (p * H) * G[4];
    double l1 = vel[d] - a, l2 = vel[d], l3 = vel[d] + a;
    if (nsg[d] * l1 > 0.0) l1 = 0.0;
    if (nsg[d] * l2 > 0.0) l2 = 0.0;
    if (nsg[d] * l3 > 0.0) l3 = 0.0;
    const double dun = dq[1 + d];
    const double La = 0.5 * l1 * (ovra * dq[4] - dun);
    const double Lb = 0.5 * l3 * (ovra * dq[4] + dun);
    const double Le = l2 * (dq[0] - ova2 * dq[4]);
    double P[5];
    P[0] = rova * (Lb + La) + Le;
    P[4] = rhoa * (Lb + La);
#pragma unroll
    for (int t = 0; t < 3; ++t) P[1 + t] = (t == d) ? (Lb - La) : l2 * dq[1 + t];
    const double wd = wgt[d];
    out[0] += wd * P[0];
    out[1] += wd * (u * P[0] + rho * P[1]);
    out[2] += wd * (v * P[0] + rho * P[2]);
    out[3] += wd * (w * P[0] + rho * P[3]);
    out[4] += wd * (ek * P[0] + rhou * P[1] + rhov * P[2] + rhow * P[3] + P[4] / kGm1);
  }
}

// ========================================================== volume kernel
// One workgroup of NT threads per element. LDS holds the modal state, the
// sum-factorisation intermediates and the primitive state at the points;
// the 1D operators come from the kernel arguments (scalar registers) in the
// pencil stages and from LDS rows in the per-point stages.
// Workgroup size of the volume kernel: 256 threads, or 512 when the two point
// passes or the DOFs do not fit 256 threads (K = 7, M = 9).
constexpr int vol_threads(int K, int M)
{
  return K * K * K <= 256 && M * M * M <= 256 + (256 / M) * M ? 256 : 512;
}

// LDS layout of the volume kernel and the waves per SIMD it is compiled for
// (NT = 256: three workgroups per CU when the LDS fits three times into 160 KB,
// which caps the kernel at 168 VGPRs; NT = 512: one or more workgroups of
// two waves per SIMD, or two workgroups when the LDS allows, which caps the
// kernel at 128 VGPRs; NT = 64, 128: as many workgroups as the LDS allows, at
// most three waves per SIMD).
template <int K, int M, int NT, typename T = double>
struct VolShape {
  static constexpr int ND = K * K * K, NI = M * M * M, NR = M * M;
  // Points go through the pointwise stages in two passes split at a row
  // (jk) boundary: rows [0, R0) in the first, the remaining rows in the second.
  static constexpr int R0 = (NT / M) < NR ? (NT / M) : NR;
  static constexpr int P0 = R0 * M, P1 = NI - P0;
  static constexpr int SZ = 5 * M * K * K;          // one z-stage buffer
  static constexpr int SY = 15 * NR * K;            // y-stage buffers, [jk][3][5][K]
  static constexpr int SF0 = 15 * P0, SF = 15 * NI; // fluxes of the first pass / all
  static constexpr int SI = 15 * NR * K;            // backward i-stage output
  static constexpr int BIG = cmax(cmax(2 * SZ + SY, SF), SI + 10 * M * K * K);
  // variable stride of the modal state in LDS: congruent to K^2 modulo the
  // bank count in elements, so that the (variable, j, i) pencil items of a
  // wave run on across a variable boundary without bank conflicts
  static constexpr int MODB = 256 / (int) sizeof(T);
  static constexpr int SUV = ND + ((K * K - ND) % MODB + MODB) % MODB;
  static constexpr int LDS = (int) sizeof(T) * (2 * M * K + 4 * SUV + ND + BIG + 15) + 8 * (45 + 6 * (NT / 64) + 4 * K + 5);
  static constexpr int WGS = (160 * 1024) / LDS;   // workgroups per CU by LDS
  // FP32 contractions: the smaller LDS allows more waves per SIMD
#ifndef VCP_MIXED_CG_TOL
#define VCP_MIXED_CG_TOL 1e-5
#endif
#ifndef VCP_F32_WAVES
#define VCP_F32_WAVES 3
#endif
  static constexpr int WAVES = NT == 256 ? (sizeof(T) == 4 && WGS >= VCP_F32_WAVES ? VCP_F32_WAVES : WGS >= 3 ? 3 : 2)
                              : NT > 256 ? (NT / 256) * (WGS >= 2 ? 2 : 1)
                                         : cmax(1, WGS * (NT / 64) / 4 < 3 ? WGS * (NT / 64) / 4 : 3);
  static_assert(WGS >= 1, "volume kernel LDS exceeds 160 KB");
  // mass-matrix solve buffers cz, cy, pI, by inside sBig
  static_assert(SZ + 5 * M * M * K + 5 * NI + 5 * M * K * K <= BIG, "CG buffers");
  // A p in the padded state layout runs from cz into the (then dead) cy
  static_assert(4 * SUV + ND <= SZ + 5 * M * M * K, "A p buffer");
};

// blockIdx.x is the element, or (BATCH) a work item (case, element) of the
// group; a is the case part. T is the type of the sum-factorisation
// contractions, the LDS buffers and the mass-matrix solve: double, or float
// for the mixed-precision form, where the state, its element means, the
// pointwise physics, the face-slot gather, the metric and the RK update stay
// in double. In the float form the mean (mode 0) of each variable is taken
// out of the modal state before the contractions and added back at the
// points in double (it contributes the same constant at every point and
// nothing to the gradients), so the float contractions carry only the
// variation about the mean.
template <int K, int M, int NT, bool BATCH, typename T = double>
__global__ __launch_bounds__(NT, (VolShape<K, M, NT, T>::WAVES)) void volume_kernel(VolArgs<K, M, T> args)
{
  using S = VolShape<K, M, NT, T>;
  constexpr bool kMixed = sizeof(T) == 4;
  constexpr int ND = S::ND, NI = S::NI, NR = S::NR, R0 = S::R0, P0 = S::P0, P1 = S::P1;
  constexpr int SZ = S::SZ, SF0 = S::SF0, SI = S::SI, BIG = S::BIG;
  static_assert(P1 <= NT, "two point passes");
  // the second point pass runs on the last waves (from thread T2), which own
  // no DOFs, leaving the first waves free
  constexpr int T2 = NT - ((P1 + 63) / 64) * 64;
  constexpr bool kPrecondLate = P1 > 0 && T2 >= 64;
  // rot() below wraps thread numbers with & (NT - 1)
  static_assert((NT & (NT - 1)) == 0, "NT must be a power of two");
  // the first pass's fluxes must leave the y-stage rows of the second pass intact
  static_assert(P1 == 0 || SF0 <= 2 * SZ + 15 * R0 * K, "flux / y-stage overlap");
  __shared__ T sA[M * K], sD[M * K];
  constexpr int SUV = S::SUV;
  __shared__ T sU[4 * SUV + ND];
  __shared__ T sBig[BIG];
  __shared__ double sCm[5][9];
  __shared__ double sRed[2 * (NT / 64) * 3];
  __shared__ T sDv[15];
  __shared__ double sFv[4 * K];
  __shared__ double sMean[5];     // float form: mode 0 of each variable
  // the case part is read where it is used: from the kernel arguments, or
  // (BATCH) from the group's case table
  const CaseDev *ap;
  int e;
  if constexpr (BATCH) {
    const WorkItem w = args.work[args.w0 + blockIdx.x];
    ap = args.cases + w.c;
    e = w.e;
    if (ap->flags[kWordActive] == 0) return;
  } else {
    ap = &args.c;
    e = blockIdx.x;
  }
  const CaseDev &a = *ap;
  const int tid = threadIdx.x;
  // Operators of the pencil stages: the kernel arguments (scalar registers)
  // while they fit; beyond that the compiler parks them in VGPR lanes and
  // reads them back with v_readlane (28 % of the instructions at K = 7 in
  // FP64), so larger operators come from the LDS copies: both above 800 bytes
  // (K = 7 in FP64: 26 % faster), D alone above 600 (K = 6 in FP64: 3 %
  // faster; both from LDS there, or either at K <= 5, measured slower).
  constexpr bool kArgOpsA = sizeof(args.op) <= 800, kArgOpsD = sizeof(args.op) <= 600;
  const T *A = kArgOpsA ? args.op.A : sA, *D = kArgOpsD ? args.op.D : sD;
  const double *Ue = a.U + (size_t) e * 5 * ND;
  const double *vg = a.vgeo + (size_t) e * kVolGeo * NI;
  const int sp = a.spIdx ? a.spIdx[e] : -1;
  for (int i = tid; i < M * K; i += NT) { sA[i] = args.op.A[i]; sD[i] = args.op.D[i]; }
  for (int i = tid; i < 4 * K; i += NT) sFv[i] = args.fvec[i / K][i % K];
  {
    // the modal state: every load issued before the first is waited for
    constexpr int NIT = (5 * ND + NT - 1) / NT;
    double uu[NIT];
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
      const int i = tid + it * NT;
      uu[it] = Ue[i < 5 * ND ? i : 0];
    }
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
      const int i = tid + it * NT;
      if (i < 5 * ND) {
        if constexpr (kMixed) {
          if (i % ND == 0) sMean[i / ND] = uu[it];
          sU[(i / ND) * SUV + i % ND] = i % ND == 0 ? T(0) : T(uu[it]);
        } else {
          sU[(i / ND) * SUV + i % ND] = uu[it];
        }
      }
    }
  }
  if (args.bug)
    for (int i = tid; i < 45; i += NT) sCm[i / 9][i % 9] = vg[(1 + i % 9) * NI + i / 9];
  // loaded here, used much later, so that their latency passes under the
  // first contractions: the element's Jacobian constant and (K <= 5, where
  // the registers allow it) the metric of this thread's point in the first
  // point pass
  const double elemJac = a.econst[(size_t) e * kElemConst + 3];
  constexpr bool kGeoEarly = K <= 5;
  double geoJ0 = 0.0, geoW0 = 0.0, geoCm0[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (kGeoEarly && tid < P0) {
    geoJ0 = vg[tid];
    geoW0 = a.tables[args.to.w3 + tid];
#pragma unroll
    for (int t = 0; t < 9; ++t) geoCm0[t] = vg[(1 + t) * NI + tid];
  }
  __syncthreads();

  // ---------------- step 1: solution and parametric gradients at the points
  // z-stage buffers, then the y-stage buffers row-major: yB[jk][3][5][K] with
  // the (A, A), (D, A), (A, D) contractions of variable v at row jk = k * M + j
  T *zA = sBig, *zD = zA + SZ, *yB = zD + SZ;
  for (int pc = tid; pc < 5 * K * K; pc += NT) {      // pencil (v, j, i) along k
    const int v = pc / (K * K), ji = pc - v * K * K;
    T x[K];
#pragma unroll
    for (int q = 0; q < K; ++q) x[q] = sU[v * SUV + q * K * K + ji];
#pragma unroll
    for (int n = 0; n < M; ++n) {
      T sa = 0, sd = 0;
#pragma unroll
      for (int q = 0; q < K; ++q) { sa += A[n * K + q] * x[q]; sd += D[n * K + q] * x[q]; }
      zA[(v * M + n) * K * K + ji] = sa;
      zD[(v * M + n) * K * K + ji] = sd;
    }
  }
  // element-centre entropy variables -> preconditioner dV/dU / Jbar in sDv,
  // by one whole wave (lane 0 the last part), in a stage where that wave has
  // no other work: the second point pass when it leaves the first wave idle,
  // else the first contraction, beside the pencils (none in the last wave when
  // 5 K^2 <= NT - 64)
  auto precond = [&]() {
    const int lane = tid & 63;
    double Vc[5] = {0, 0, 0, 0, 0};
    for (int i = lane; i < ND; i += 64) {
      const int ii = i % K, jj = (i / K) % K, kk = i / (K * K);
      double b = 0.0, bj = 0.0, bk = 0.0;
#pragma unroll
      for (int q = 0; q < K; ++q) {
        if (q == ii) b = args.c0[q];
        if (q == jj) bj = args.c0[q];
        if (q == kk) bk = args.c0[q];
      }
      b *= bj * bk;
#pragma unroll
      for (int v = 0; v < 5; ++v) Vc[v] += b * sU[v * SUV + i];
    }
#pragma unroll
    for (int v = 0; v < 5; ++v) Vc[v] = wave_sum(Vc[v]);
    if constexpr (kMixed) {
      const double c03 = args.c0[0] * args.c0[0] * args.c0[0];
#pragma unroll
      for (int v = 0; v < 5; ++v) Vc[v] += c03 * sMean[v];
    }
    if (lane == 0) {
      double rho, u, v, w, p, eKin;
      ent2prim(Vc, rho, u, v, w, p, eKin);
      const double JacInv = 8.0 / elemJac;
      const double pInv = 1.0 / p, a1 = kGm1 * rho * pInv * pInv, a2 = eKin * a1;
      sDv[0] = T(JacInv * (kGam / kGm1 / rho + a2 * eKin)); sDv[1] = T(-JacInv * u * a2); sDv[2] = T(-JacInv * v * a2);
      sDv[3] = T(-JacInv * w * a2); sDv[4] = T(JacInv * (a2 - pInv)); sDv[5] = T(JacInv * (a1 * u * u + pInv));
      sDv[6] = T(JacInv * a1 * u * v); sDv[7] = T(JacInv * a1 * u * w); sDv[8] = T(-JacInv * a1 * u);
      sDv[9] = T(JacInv * (a1 * v * v + pInv)); sDv[10] = T(JacInv * a1 * v * w); sDv[11] = T(-JacInv * a1 * v);
      sDv[12] = T(JacInv * (a1 * w * w + pInv)); sDv[13] = T(-JacInv * a1 * w); sDv[14] = T(JacInv * a1);
    }
  };
  if (!kPrecondLate && tid >= NT - 64) precond();
  __syncthreads();
  for (int pc = tid; pc < 5 * M * K; pc += NT) {      // pencil (v, k', i) along j
    const int vk = pc / K, i = pc - vk * K;
    const int v = vk / M, kp = vk - v * M;
    T xa[K], xd[K];
#pragma unroll
    for (int q = 0; q < K; ++q) { xa[q] = zA[(vk * K + q) * K + i]; xd[q] = zD[(vk * K + q) * K + i]; }
    T *yo = yB + (kp * M * 15 + v) * K + i;
#pragma unroll
    for (int n = 0; n < M; ++n) {
      T s0 = 0, s1 = 0, s2 = 0;
#pragma unroll
      for (int q = 0; q < K; ++q) { s0 += A[n * K + q] * xa[q]; s1 += D[n * K + q] * xa[q]; s2 += A[n * K + q] * xd[q]; }
      yo[(n * 15 + 0) * K] = s0;
      yo[(n * 15 + 5) * K] = s1;
      yo[(n * 15 + 10) * K] = s2;
    }
  }
  __syncthreads();

  // ---------------- steps 2, 3, 5, 6: x stage, pointwise fluxes, primitive
  // state, monitoring; in two passes over the points (rows [0, R0), then the
  // rest), each thread one point per pass. The first pass's fluxes F[b][m]
  // (b = 5 d + l) fill [0, SF0) and leave the y-stage rows of the second
  // pass in place; the second pass's fluxes go to SF0 + [b][m - P0].
  const double factHF = a.mu * kGam / kPrLam;
  const double lenLES = a.sgs ? a.econst[(size_t) e * kElemConst + kElemLenLES] : 0.0;
  double mach2 = 0.0, evMax = 0.0;
  // (w J rho, u, v, w, w J p, w J dU/dV_44) at the thread's points, for the
  // weighted dU/dV in the solve
  T prim[2][6];
  // float form: the constant contribution of the element means at every point
  double meanPt[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  if constexpr (kMixed) {
    const double c03 = args.c0[0] * args.c0[0] * args.c0[0];
#pragma unroll
    for (int v = 0; v < 5; ++v) meanPt[v] = c03 * sMean[v];
  }
#pragma unroll
  for (int q = 0; q < (P1 > 0 ? 2 : 1); ++q) {
    const int m = q == 0 ? tid : P0 + tid - T2;
    const bool act = q == 0 ? tid < P0 : (tid >= T2 && tid - T2 < P1);
    double Vp[5], Gr[5], Gs[5], Gt[5];
    // the second pass's metric, loaded before its x stage
    double geoJ1 = 0.0, geoW1 = 0.0, geoCm1[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (kGeoEarly && q == 1 && act) {
      geoJ1 = vg[m];
      geoW1 = a.tables[args.to.w3 + m];
#pragma unroll
      for (int t = 0; t < 9; ++t) geoCm1[t] = vg[(1 + t) * NI + m];
    }
    if (act) {
      const int i = m % M, jk = m / M;
      T ai[K], di[K];
#pragma unroll
      for (int ii = 0; ii < K; ++ii) { ai[ii] = sA[i * K + ii]; di[ii] = sD[i * K + ii]; }
      int roff = jk * 15 * K;
      asm volatile("" : "+v"(roff));   // one base register, immediate LDS offsets
      const T *row = yB + roff;
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        const T *ra = row + v * K, *rd = row + (5 + v) * K, *rz = row + (10 + v) * K;
        T s0 = 0, s1 = 0, s2 = 0, s3 = 0;
#pragma unroll
        for (int ii = 0; ii < K; ++ii) {
          const T x = ra[ii];
          s0 += ai[ii] * x; s1 += di[ii] * x; s2 += ai[ii] * rd[ii]; s3 += ai[ii] * rz[ii];
        }
        Vp[v] = kMixed ? double(s0) + meanPt[v] : double(s0);
        Gr[v] = s1; Gs[v] = s2; Gt[v] = s3;
      }
    }
    __syncthreads();
    if (kPrecondLate && q == 1 && tid < 64) precond();
    if (!act) continue;
    T *Fb = q == 0 ? sBig + m : sBig + SF0 + (m - P0);
    const int fstride = q == 0 ? P0 : P1;
    const double J = !kGeoEarly ? vg[m] : q == 0 ? geoJ0 : geoJ1;
    const double wJ = (!kGeoEarly ? a.tables[args.to.w3 + m] : q == 0 ? geoW0 : geoW1) * J;
    double cm[9];                                      // dr/dx .. dt/dz at the point
#pragma unroll
    for (int t = 0; t < 9; ++t) cm[t] = !kGeoEarly ? vg[(1 + t) * NI + m] : q == 0 ? geoCm0[t] : geoCm1[t];
    double gx[5], gy[5], gz[5];
    if (args.bug) {
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        const double *c = sCm[v];
        gx[v] = Gr[v] * c[0] + Gs[v] * c[3] + Gt[v] * c[6];
        gy[v] = Gr[v] * c[1] + Gs[v] * c[4] + Gt[v] * c[7];
        gz[v] = Gr[v] * c[2] + Gs[v] * c[5] + Gt[v] * c[8];
      }
    } else {
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        gx[v] = Gr[v] * cm[0] + Gs[v] * cm[3] + Gt[v] * cm[6];
        gy[v] = Gr[v] * cm[1] + Gs[v] * cm[4] + Gt[v] * cm[7];
        gz[v] = Gr[v] * cm[2] + Gs[v] * cm[5] + Gt[v] * cm[8];
      }
    }
    double rho, u, v, w, p, eKin;
    ent2prim(Vp, rho, u, v, w, p, eKin);
    const double V4Inv = 1.0 / Vp[4];
    const double rE = p * kInvGm1 + rho * eKin;
    prim[q][0] = T(wJ * rho); prim[q][1] = T(u); prim[q][2] = T(v); prim[q][3] = T(w); prim[q][4] = T(wJ * p);
    prim[q][5] = T(wJ * (rE * (rE + p) / rho + p * eKin));
    const double dudx = -V4Inv * (gx[1] + u * gx[4]), dudy = -V4Inv * (gy[1] + u * gy[4]), dudz = -V4Inv * (gz[1] + u * gz[4]);
    const double dvdx = -V4Inv * (gx[2] + v * gx[4]), dvdy = -V4Inv * (gy[2] + v * gy[4]), dvdz = -V4Inv * (gz[2] + v * gz[4]);
    const double dwdx = -V4Inv * (gx[3] + w * gx[4]), dwdy = -V4Inv * (gy[3] + w * gy[4]), dwdz = -V4Inv * (gz[3] + w * gz[4]);
    const double fe = V4Inv * V4Inv * kInvGm1;
    const double dedx = fe * gx[4], dedy = fe * gy[4], dedz = fe * gz[4];
    if (args.ctl.monitor) mach2 = fmax(mach2, rho * (u * u + v * v + w * w) / (kGam * p));
    const double ru = rho * u, rv = rho * v, rw = rho * w;
    // eddy viscosity of the SGS model (ElementClass.cpp:4268-4287, :4305-4306)
    // in the total viscosity and heat-flux factor (:4339-4342)
    double mu = a.mu, fhT = factHF;
    if (a.sgs != kSgsNone) {
      const double G[3][3] = {{dudx, dudy, dudz}, {dvdx, dvdy, dvdz}, {dwdx, dwdy, dwdz}};
      const double ev = eddy_viscosity(a.sgs, lenLES, rho, G);
      evMax = fmax(evMax, ev);
      mu = a.mu + ev;
      fhT = a.mu * (kGam / kPrLam) + ev * (kGam / kPrTurb);
    }
    const double divV = kLamOverMu * (dudx + dvdy + dwdz);
    const double txx = mu * (2.0 * dudx + divV), tyy = mu * (2.0 * dvdy + divV), tzz = mu * (2.0 * dwdz + divV);
    const double txy = mu * (dudy + dvdx), txz = mu * (dudz + dwdx), tyz = mu * (dvdz + dwdy);
    const double qx = -fhT * dedx, qy = -fhT * dedy, qz = -fhT * dedz;
    const double fx[5] = {ru, ru * u - txx + p, rv * u - txy, rw * u - txz, (rE + p) * u - (u * txx + v * txy + w * txz - qx)};
    const double fy[5] = {rv, ru * v - txy, rv * v - tyy + p, rw * v - tyz, (rE + p) * v - (u * txy + v * tyy + w * tyz - qy)};
    const double fz[5] = {rw, ru * w - txz, rv * w - tyz, rw * w - tzz + p, (rE + p) * w - (u * txz + v * tyz + w * tzz - qz)};
    const double weight = -wJ;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double mx = weight * cm[3 * d], my = weight * cm[3 * d + 1], mz = weight * cm[3 * d + 2];
#pragma unroll
      for (int l = 0; l < 5; ++l) Fb[(d * 5 + l) * fstride] = T(fx[l] * mx + fy[l] * my + fz[l] * mz);
    }
    if (sp >= 0) {
      // sponge damping sigma J w (U - U_target) and CML sources
      // (ElementClass.cpp:4450-4489, :3887-4117)
      const double *fn = a.dampFn + (size_t) sp * 3 * NI, *ds = a.dampState + (size_t) sp * 5 * NI;
      const double sigma = wJ * (fn[m] + fn[NI + m] + fn[2 * NI + m]);
      double src[5];
      src[0] = sigma * (rho - ds[m]);
      src[1] = sigma * (ru - ds[NI + m]);
      src[2] = sigma * (rv - ds[2 * NI + m]);
      src[3] = sigma * (rw - ds[3 * NI + m]);
      src[4] = sigma * (rE - ds[4 * NI + m]);
      if (a.cml) {
        const double *cf = a.spCoef + (size_t) sp * 8;
        const double wgt[3] = {weight * cf[0] * fn[m], weight * cf[1] * fn[NI + m], weight * cf[2] * fn[2 * NI + m]};
        const double nsg[3] = {cf[3], cf[4], cf[5]};
        cml_source(Vp, gx, gy, gz, wgt, nsg, src);
      }
      double *sb = a.srcBuf + (size_t) sp * 5 * NI;
#pragma unroll
      for (int l = 0; l < 5; ++l) sb[l * NI + m] = src[l];
    }

  }
  __syncthreads();

  if (a.bfBuf) {
    // body force per unit volume (ElementClass.cpp:4492-4523): -w J f on the
    // momentum and -w J u.f on the energy equation at this thread's points,
    // with the velocities kept for the solve; projected with the sponge source
    const double fx = a.fbody[0], fy = a.fbody[1], fz = a.fbody[2];
    double *bb = a.bfBuf + (size_t) e * 4 * NI;
#pragma unroll
    for (int q = 0; q < (P1 > 0 ? 2 : 1); ++q) {
      const int m = q == 0 ? tid : P0 + tid - T2;
      if (!(q == 0 ? tid < P0 : (tid >= T2 && tid - T2 < P1))) continue;
      const double weight = -(a.tables[args.to.w3 + m] * vg[m]);
      const double u = prim[q][1], v = prim[q][2], w = prim[q][3];
      bb[m] = weight * fx;
      bb[NI + m] = weight * fy;
      bb[2 * NI + m] = weight * fz;
      bb[3 * NI + m] = weight * (u * fx + v * fy + w * fz);
    }
  }

  // the six face contributions at this thread's DOF (DOF-major: thread
  // d < ND owns DOF d of all five variables): each face slot holds the value
  // and normal-derivative coefficient planes over its two tangential axes,
  // times the face vectors phi(r_f), phi'(r_f) at the DOF's normal index
  static_assert(ND <= NT, "one DOF per thread in the mass-matrix solve");
  double racc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  if (tid < ND) {
    const int di = tid % K, dj = (tid / K) % K, dk = tid / (K * K);
    const double *fr = a.faceRes + (size_t) e * 6 * 10 * K * K;
#pragma unroll
    for (int f = 0; f < 6; ++f) {
      const int dir = f >> 1, side = f & 1;
      const int tn = dir == 0 ? di : dir == 1 ? dj : dk;
      const int fa = dir == 0 ? dj : di;
      const int sl = dir == 2 ? dj : dk;
      const double cv = sFv[side * K + tn], cd = sFv[(2 + side) * K + tn];
      const double *b = fr + f * 10 * K * K + sl * K + fa;
#pragma unroll
      for (int l = 0; l < 5; ++l) racc[l] += cv * b[l * K * K] + cd * b[(5 + l) * K * K];
    }
  }

  // ---------------- step 7: transposed contractions of the three fluxes
  // rows (b, jk) along i, compacted to I[b][jk][K] once every row is read
  {
    constexpr int NRW = 15 * NR, RPT = (NRW + NT - 1) / NT;
    T o[RPT][K];
#pragma unroll
    for (int t = 0; t < RPT; ++t) {
      const int r = tid + t * NT;
      if (r < NRW) {
        const int b = r / NR, jk = r - b * NR;
        const T *x = jk < R0 ? sBig + b * P0 + jk * M : sBig + SF0 + b * P1 + (jk - R0) * M;
        T xx[M];
#pragma unroll
        for (int q = 0; q < M; ++q) xx[q] = x[q];
#pragma unroll
        for (int n = 0; n < K; ++n) {
          T s = 0;
          if (b < 5) {
#pragma unroll
            for (int q = 0; q < M; ++q) s += D[q * K + n] * xx[q];
          } else {
#pragma unroll
            for (int q = 0; q < M; ++q) s += A[q * K + n] * xx[q];
          }
          o[t][n] = s;
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < RPT; ++t) {
      const int r = tid + t * NT;
      if (r < NRW)
#pragma unroll
        for (int n = 0; n < K; ++n) sBig[r * K + n] = o[t][n];
    }
  }
  __syncthreads();
  const T *I0 = sBig, *I1 = I0 + 5 * NR * K, *I2 = I1 + 5 * NR * K;
  T *Y01 = sBig + SI, *Y2 = Y01 + 5 * M * K * K;   // after the i-stage block
  for (int pc = tid; pc < 5 * M * K; pc += NT) {      // pencil (v, k, i') along j
    const int vk = pc / K, i = pc - vk * K;
    T x0[M], x1[M], x2[M];
#pragma unroll
    for (int q = 0; q < M; ++q) {
      x0[q] = I0[(vk * M + q) * K + i]; x1[q] = I1[(vk * M + q) * K + i]; x2[q] = I2[(vk * M + q) * K + i];
    }
#pragma unroll
    for (int n = 0; n < K; ++n) {
      T s01 = 0, s2 = 0;
#pragma unroll
      for (int q = 0; q < M; ++q) { s01 += A[q * K + n] * x0[q] + D[q * K + n] * x1[q]; s2 += A[q * K + n] * x2[q]; }
      Y01[(vk * K + n) * K + i] = s01;
      Y2[(vk * K + n) * K + i] = s2;
    }
  }
  __syncthreads();
  for (int pc = tid; pc < 5 * K * K; pc += NT) {      // pencil (v, j', i') along k
    const int v = pc / (K * K), ji = pc - v * K * K;
    T y0[M], y2[M];
#pragma unroll
    for (int q = 0; q < M; ++q) { y0[q] = Y01[(v * M + q) * K * K + ji]; y2[q] = Y2[(v * M + q) * K * K + ji]; }
#pragma unroll
    for (int n = 0; n < K; ++n) {
      T s = 0;
#pragma unroll
      for (int q = 0; q < M; ++q) s += A[q * K + n] * y0[q] + D[q * K + n] * y2[q];
      sU[v * SUV + n * K * K + ji] = s;
    }
  }
  __syncthreads();
  if (sp >= 0 || a.bfBuf) {
    // the sponge element's damping and CML sources and the body force
    // (ElementClass.cpp:4492-4533, on the momentum and energy equations),
    // both from the point pass, summed at the points and projected with the
    // basis functions: rows along i, pencils along j, then along k,
    // accumulated into the residual
    const double *sb = sp >= 0 ? a.srcBuf + (size_t) sp * 5 * NI : nullptr;
    const double *bb = a.bfBuf ? a.bfBuf + (size_t) e * 4 * NI : nullptr;
    for (int r = tid; r < 5 * M * M; r += NT) {       // rows along i
      T xx[M];
#pragma unroll
      for (int q = 0; q < M; ++q) {
        double x = sb ? sb[r * M + q] : 0.0;
        if (bb && r >= M * M) x += bb[r * M + q - NI];
        xx[q] = T(x);
      }
#pragma unroll
      for (int n = 0; n < K; ++n) {
        T s = 0;
#pragma unroll
        for (int q = 0; q < M; ++q) s += A[q * K + n] * xx[q];
        sBig[r * K + n] = s;
      }
    }
    __syncthreads();
    T *Ys = sBig + 5 * M * M * K;
    for (int pc = tid; pc < 5 * M * K; pc += NT) {    // pencils along j
      const int vk = pc / K, i = pc - vk * K;
      T xx[M];
#pragma unroll
      for (int q = 0; q < M; ++q) xx[q] = sBig[(vk * M + q) * K + i];
#pragma unroll
      for (int n = 0; n < K; ++n) {
        T s = 0;
#pragma unroll
        for (int q = 0; q < M; ++q) s += A[q * K + n] * xx[q];
        Ys[(vk * K + n) * K + i] = s;
      }
    }
    __syncthreads();
    for (int pc = tid; pc < 5 * K * K; pc += NT) {    // pencils along k, accumulated
      const int v = pc / (K * K), ji = pc - v * K * K;
      T xx[M];
#pragma unroll
      for (int q = 0; q < M; ++q) xx[q] = Ys[(v * M + q) * K * K + ji];
#pragma unroll
      for (int n = 0; n < K; ++n) {
        T s = 0;
#pragma unroll
        for (int q = 0; q < M; ++q) s += A[q * K + n] * xx[q];
        sU[v * SUV + n * K * K + ji] += s;
      }
    }
    __syncthreads();
  }

  // ---------------- gather the six face contributions
  // (a non-finite right-hand side is flagged: the solve's max-norms would
  // pass over it)
  bool bad = false;
  T r[5];
  if constexpr (kMixed) {
#pragma unroll
    for (int l = 0; l < 5; ++l) {
      if (tid < ND) racc[l] += double(sU[l * SUV + tid]);
      bad = bad || !isfinite(racc[l]);
      r[l] = T(racc[l]);
    }
  } else {
    if (tid < ND)
#pragma unroll
      for (int l = 0; l < 5; ++l) {
        racc[l] += sU[l * SUV + tid];
        bad = bad || !isfinite(racc[l]);
      }
#pragma unroll
    for (int l = 0; l < 5; ++l) r[l] = racc[l];
  }

  // ---------------- entropy-variable mass matrix: flexible PCG
  const double JacInv = 8.0 / elemJac;
  const bool own = tid < ND;
  T x[5], z[5], pv[5];
  int half = 0;
  double r0 = 0.0;
#pragma unroll
  for (int l = 0; l < 5; ++l) { x[l] = 0; r0 = fmax(r0, fabs(double(r[l]))); }
  {
    // with the element's maximum Mach^2 and eddy viscosity for the
    // monitoring data
    double red[3] = {r0, mach2, evMax};
    block_reduce1<NT, 3, 0>(red, sRed, half);
    r0 = red[0];
    if (args.ctl.monitor && tid == 0) a.mach2Elem[e] = red[1];
    if (a.eddyPart && tid == 0) a.eddyPart[e] = red[2];
  }
  int iters = 0;
  if (r0 * JacInv > 1e-10) {
    T *cz = sBig, *cy = cz + SZ, *pI = cy + 5 * M * M * K, *by = pI + 5 * NI;
    // z = dV/dU(centre) r, local to the thread owning the DOF
#define VCP_PRECOND()                                                                             \
    do {                                                                                          \
      z[0] = sDv[0] * r[0] + sDv[1] * r[1] + sDv[2] * r[2] + sDv[3] * r[3] + sDv[4] * r[4];       \
      z[1] = sDv[1] * r[0] + sDv[5] * r[1] + sDv[6] * r[2] + sDv[7] * r[3] + sDv[8] * r[4];       \
      z[2] = sDv[2] * r[0] + sDv[6] * r[1] + sDv[9] * r[2] + sDv[10] * r[3] + sDv[11] * r[4];     \
      z[3] = sDv[3] * r[0] + sDv[7] * r[1] + sDv[10] * r[2] + sDv[12] * r[3] + sDv[13] * r[4];    \
      z[4] = sDv[4] * r[0] + sDv[8] * r[1] + sDv[11] * r[2] + sDv[13] * r[3] + sDv[14] * r[4];    \
    } while (0)
    VCP_PRECOND();
    double rz = 0.0;
#pragma unroll
    for (int l = 0; l < 5; ++l) { pv[l] = z[l]; rz += double(r[l]) * double(z[l]); }
    { double red[1] = {rz}; block_reduce1<NT, 1, 1, (ND + 63) / 64>(red, sRed, half); rz = red[0]; }
    // Stages with fewer items than threads start at different waves, so that
    // the waves of the workgroup (one per SIMD) carry similar shares of the
    // iteration: the DOF owners and the second point pass sit in the first
    // waves, the two k stages in the last ones. rot(F) numbers the threads
    // from wave F on.
    constexpr int NWV = NT / 64, NWK = (5 * K * K + 63) / 64;
    constexpr int FK = NWV - NWK < 2 ? NWV - NWK : 2;
    auto rot = [tid](int F) { return (tid + NT - 64 * F) & (NT - 1); };
    for (int it = 0;; ++it) {
      if (it == 100) { atomicOr(a.flags, kFlagCGNotConverged); break; }
      if (own)
#pragma unroll
        for (int l = 0; l < 5; ++l) sU[l * SUV + tid] = pv[l];
      __syncthreads();
      for (int pc = rot(FK); pc < 5 * K * K; pc += NT) {
        const int v = pc / (K * K), ji = pc - v * K * K;
        T xx[K];
#pragma unroll
        for (int q = 0; q < K; ++q) xx[q] = sU[v * SUV + q * K * K + ji];
#pragma unroll
        for (int n = 0; n < M; ++n) {
          T s = 0;
#pragma unroll
          for (int q = 0; q < K; ++q) s += A[n * K + q] * xx[q];
          cz[(v * M + n) * K * K + ji] = s;
        }
      }
      __syncthreads();
      for (int pc = rot(1); pc < 5 * M * K; pc += NT) {
        const int i = pc / (5 * M), vk = pc - i * (5 * M);
        T xx[K];
#pragma unroll
        for (int q = 0; q < K; ++q) xx[q] = cz[(vk * K + q) * K + i];
#pragma unroll
        for (int n = 0; n < M; ++n) {
          T s = 0;
#pragma unroll
          for (int q = 0; q < K; ++q) s += A[n * K + q] * xx[q];
          cy[(vk * M + n) * K + i] = s;
        }
      }
      __syncthreads();
      // along i, one pencil (variable, row jk) per item with the operator from
      // the kernel arguments: K LDS reads per M outputs, where a per-point
      // contraction would read 5 K values and an operator row per point
      for (int pc = tid; pc < 5 * M * M; pc += NT) {
        const int l = pc / (M * M), jk = pc - l * M * M;
        T xx[K];
#pragma unroll
        for (int q = 0; q < K; ++q) xx[q] = cy[pc * K + q];
#pragma unroll
        for (int n = 0; n < M; ++n) {
          T s = 0;
#pragma unroll
          for (int q = 0; q < K; ++q) s += A[n * K + q] * xx[q];
          pI[l * NI + jk * M + n] = s;
        }
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < (P1 > 0 ? 2 : 1); ++q) {
        const int m = q == 0 ? tid : P0 + tid - T2;
        if (!(q == 0 ? tid < P0 : (tid >= T2 && tid - T2 < P1))) continue;
        T pp[5];
#pragma unroll
        for (int l = 0; l < 5; ++l) pp[l] = pI[l * NI + m];
        // w J dU/dV at the point (ElementClass.cpp:4228-4253): every entry is
        // homogeneous of degree one in (rho, p), so the scaled pair held in
        // prim gives the weighted matrix directly
        const T rho = prim[q][0], u = prim[q][1], v = prim[q][2], w = prim[q][3], p = prim[q][4];
        const T d44 = prim[q][5];
        const T eKin = T(0.5) * (u * u + v * v + w * w);
        const T rE = p * T(kInvGm1) + rho * eKin, rH = rE + p;
        const T ru = rho * u, rv = rho * v, rw = rho * w;
        pI[0 * NI + m] = rho * pp[0] + ru * pp[1] + rv * pp[2] + rw * pp[3] + rE * pp[4];
        pI[1 * NI + m] = ru * pp[0] + (ru * u + p) * pp[1] + ru * v * pp[2] + ru * w * pp[3] + rH * u * pp[4];
        pI[2 * NI + m] = rv * pp[0] + ru * v * pp[1] + (rv * v + p) * pp[2] + rv * w * pp[3] + rH * v * pp[4];
        pI[3 * NI + m] = rw * pp[0] + ru * w * pp[1] + rv * w * pp[2] + (rw * w + p) * pp[3] + rH * w * pp[4];
        pI[4 * NI + m] = rE * pp[0] + rH * u * pp[1] + rH * v * pp[2] + rH * w * pp[3] + d44 * pp[4];
      }
      __syncthreads();
      for (int rr = tid; rr < 5 * M * M; rr += NT) {
        T xx[M];
#pragma unroll
        for (int q = 0; q < M; ++q) xx[q] = pI[rr * M + q];
#pragma unroll
        for (int n = 0; n < K; ++n) {
          T s = 0;
#pragma unroll
          for (int q = 0; q < M; ++q) s += A[q * K + n] * xx[q];
          pI[rr * M + n] = s;
        }
      }
      __syncthreads();
      for (int pc = rot(2); pc < 5 * M * K; pc += NT) {
        const int i = pc / (5 * M), vk = pc - i * (5 * M);
        T xx[M];
#pragma unroll
        for (int q = 0; q < M; ++q) xx[q] = pI[(vk * M + q) * M + i];
#pragma unroll
        for (int n = 0; n < K; ++n) {
          T s = 0;
#pragma unroll
          for (int q = 0; q < M; ++q) s += A[q * K + n] * xx[q];
          by[(vk * K + n) * K + i] = s;
        }
      }
      __syncthreads();
      // Ap along k, with the partial p.Ap of the entries each thread computes
      double pAp = 0.0;
      for (int pc = rot(FK); pc < 5 * K * K; pc += NT) {
        const int v = pc / (K * K), ji = pc - v * K * K;
        T xx[M];
#pragma unroll
        for (int q = 0; q < M; ++q) xx[q] = by[(v * M + q) * K * K + ji];
#pragma unroll
        for (int n = 0; n < K; ++n) {
          T s = 0;
#pragma unroll
          for (int q = 0; q < M; ++q) s += A[q * K + n] * xx[q];
          const int idx = v * SUV + n * K * K + ji;
          cz[idx] = s;
          pAp += double(sU[idx]) * double(s);
        }
      }
      {
        double red[1] = {pAp};
        block_reduce1<NT, 1, 1, (NWK < NWV ? NWK : NWV), (NWK < NWV ? FK : 0)>(red, sRed, half);
        pAp = red[0];
      }
      // vector updates only in the waves that own DOFs
      double zAp = 0.0, rzn = 0.0, rmax = 0.0, alpha = 0.0;
      if (own) {
        T ap[5];
#pragma unroll
        for (int l = 0; l < 5; ++l) ap[l] = cz[l * SUV + tid];
        alpha = rz / pAp;
        const T alphaT = T(alpha);
#pragma unroll
        for (int l = 0; l < 5; ++l) {
          x[l] += alphaT * pv[l];
          r[l] -= alphaT * ap[l];
          rmax = fmax(rmax, fabs(double(r[l])));
        }
        VCP_PRECOND();
#pragma unroll
        for (int l = 0; l < 5; ++l) { zAp += double(z[l]) * double(ap[l]); rzn += double(r[l]) * double(z[l]); }
      }
      {
        double red[3] = {zAp, rzn, rmax};
        block_reduce1<NT, 3, 2, (ND + 63) / 64>(red, sRed, half);
        zAp = red[0]; rzn = red[1]; rmax = red[2];
      }
      iters = it + 1;
      // the reference's stop rule (CG_MassMatrix.cpp:465); the float form
      // stops at 1e-5, a few units of its rounding level
      if (rmax < (kMixed ? VCP_MIXED_CG_TOL : 1e-10) * r0) break;
      if (own) {
        const T beta = T(-alpha * zAp / rz);
#pragma unroll
        for (int l = 0; l < 5; ++l) pv[l] = beta * pv[l] + z[l];
      }
      rz = rzn;
    }
#undef VCP_PRECOND
  }
  if (tid == 0) a.cgIter[e] = iters;
  // ---------------- outputs: R and/or the in-place RK stage update
  double *Re = a.R + (size_t) e * 5 * ND;
  double *Uw = a.U + (size_t) e * 5 * ND;
  const double *U0e = a.U0 + (size_t) e * 5 * ND;
  const double bdt = args.ctl.update ? args.ctl.b * a.scal[kScalDt] : 0.0;
  if (own)
#pragma unroll
    for (int l = 0; l < 5; ++l) {
      const int i = l * ND + tid;
      const double xl = double(x[l]);
      if (!isfinite(xl)) bad = true;
      if (args.ctl.store_res) Re[i] = xl;
      if (args.ctl.update) Uw[i] = (1.0 - args.ctl.a) * Uw[i] + args.ctl.a * U0e[i] - bdt * xl;
    }
  if (bad) atomicOr(a.flags, kFlagNonFinite);
}

// ================================================= open-boundary kernels
template <int K, int M>
struct BndArgs {
  const double *U, *tables, *econst, *bgeo;   // bgeo: [nBFace][13][nInt2D] own face metric
  const int *bface;
  double *machPart, *faceRes, *forcePart;
  TableOffsets to;
  int bug, monitor;             // monitor: wall forces of this residual into forcePart
  double mu, tauFact, thetaHalf;
  int sgs, nElem;
  double *eddyPart;             // eddy-viscosity maxima (null: no SGS model)
  NscbcParams nb;
};

// Operators and trace of one element face in LDS (runtime direction).
template <int K, int M>
struct BndSmem {
  double A[M * K], D[M * K], F[2][K], P1[M * M], dL[M * M], W2[M * M];
  double B[2][5 * K * K];          // normal contraction: value, normal derivative
  double T[3][5 * K * M];          // fast contraction: A, D (of value), A (of normal derivative)
  double S[4][5 * M * M];          // interior trace: value, d/d(normal), d/d(fast), d/d(slow)
  double Q[4][5 * M * M];          // boundary state: value, d/d(normal), d/d(fast), d/d(slow)
  double X[2][5 * M * M];          // B matrix / PC stage; later flux + spare
  double Sym[3][5 * M * M];        // symmetrising fluxes (normal, fast, slow)
  double red[2];
};

// Trace of element e's face (dir, side): values and parametric derivatives
// in the face-local (normal, fast, slow) slots, [5][slow][fast] layout
// (TensorProductSolAndGrad{I,J,K}Face).
template <int K, int M>
__device__ void bnd_trace(BndSmem<K, M> &sm, const double *Ue, int dir, int side, bool grads)
{
  constexpr int ND = K * K * K, N2 = M * M;
  const int tid = threadIdx.x;
  const int SN = dir == 0 ? 1 : dir == 1 ? K : K * K;
  const int SFA = dir == 0 ? K : 1;
  const int SSL = dir == 2 ? K : K * K;
  const double *Fv = sm.F[0], *dFv = sm.F[1];
  for (int pc = tid; pc < 5 * K * K; pc += 64) {
    const int v = pc / (K * K), sl = (pc / K) % K, f = pc % K;
    const double *u = Ue + v * ND + sl * SSL + f * SFA;
    double b0 = 0.0, b1 = 0.0;
#pragma unroll
    for (int t = 0; t < K; ++t) { const double x = u[t * SN]; b0 += Fv[t] * x; b1 += dFv[t] * x; }
    sm.B[0][pc] = b0; sm.B[1][pc] = b1;
  }
  __syncthreads();
  for (int pc = tid; pc < 5 * K; pc += 64) {
    double x0[K], x1[K];
#pragma unroll
    for (int f = 0; f < K; ++f) { x0[f] = sm.B[0][pc * K + f]; x1[f] = sm.B[1][pc * K + f]; }
    for (int n = 0; n < M; ++n) {
      double tA = 0.0, tD = 0.0, tN = 0.0;
#pragma unroll
      for (int f = 0; f < K; ++f) { tA += sm.A[n * K + f] * x0[f]; tD += sm.D[n * K + f] * x0[f]; tN += sm.A[n * K + f] * x1[f]; }
      sm.T[0][pc * M + n] = tA; sm.T[1][pc * M + n] = tD; sm.T[2][pc * M + n] = tN;
    }
  }
  __syncthreads();
  for (int idx = tid; idx < 5 * N2; idx += 64) {
    const int v = idx / N2, m = idx - v * N2, sp = m / M, fp = m - sp * M;
    double val = 0.0, gN = 0.0, gF = 0.0, gS = 0.0;
#pragma unroll
    for (int sl = 0; sl < K; ++sl) {
      const double a = sm.A[sp * K + sl], xa = sm.T[0][(v * K + sl) * M + fp];
      val += a * xa;
      if (grads) { gS += sm.D[sp * K + sl] * xa; gN += a * sm.T[2][(v * K + sl) * M + fp]; gF += a * sm.T[1][(v * K + sl) * M + fp]; }
    }
    sm.S[0][idx] = val;
    if (grads) { sm.S[1][idx] = gN; sm.S[2][idx] = gF; sm.S[3][idx] = gS; }
  }
  __syncthreads();
}

template <int K, int M>
__device__ void bnd_load_ops(BndSmem<K, M> &sm, const BndArgs<K, M> &a, int side)
{
  const int tid = threadIdx.x;
  const double *T = a.tables;
  for (int i = tid; i < M * K; i += 64) { sm.A[i] = T[a.to.A + i]; sm.D[i] = T[a.to.D + i]; }
  for (int i = tid; i < K; i += 64) {
    sm.F[0][i] = T[(side ? a.to.fmax : a.to.fmin) + i];
    sm.F[1][i] = T[(side ? a.to.dfmax : a.to.dfmin) + i];
  }
  for (int i = tid; i < M * M; i += 64) {
    sm.P1[i] = T[a.to.P1 + i]; sm.dL[i] = T[a.to.dLq + i]; sm.W2[i] = T[a.to.w2 + i];
  }
  __syncthreads();
}

// Weighted face-normal Mach number of one boundary face
// (WeightedMachElement, BoundaryState.cpp:935-994).
template <int K, int M>
__global__ __launch_bounds__(64) void mach_kernel(BndArgs<K, M> a)
{
  constexpr int ND = K * K * K, N2 = M * M;
  __shared__ BndSmem<K, M> sm;
  const int b = blockIdx.x;
  const int e = a.bface[2 * b], f = a.bface[2 * b + 1];
  const int dir = f / 2, side = f % 2;
  if (a.nb.bc[f] > kBCInflowTotalNSCBC) {   // not a characteristic face: no part in the average
    if (threadIdx.x == 0) a.machPart[b] = 0.0;
    return;
  }
  bnd_load_ops(sm, a, side);
  bnd_trace(sm, a.U + (size_t) e * 5 * ND, dir, side, false);
  const double *g = a.bgeo + (size_t) b * 13 * N2;
  double sum = 0.0;
  for (int m = threadIdx.x; m < N2; m += 64) {
    const double *rr = g + (size_t) (4 + 3 * dir) * N2;
    const double rx = rr[m], ry = rr[N2 + m], rz = rr[2 * N2 + m];
    const double mag = sqrt(rx * rx + ry * ry + rz * rz);
    const double V4Inv = 1.0 / sm.S[0][4 * N2 + m];
    const double u = -V4Inv * sm.S[0][N2 + m], v = -V4Inv * sm.S[0][2 * N2 + m], w = -V4Inv * sm.S[0][3 * N2 + m];
    const double c = sqrt(-kGam * V4Inv);
    const double urr = (u * rx + v * ry + w * rz) / mag;
    sum += sm.W2[m] * g[3 * N2 + m] * (urr / c);
  }
  sum = wave_sum(sum);
  if (threadIdx.x == 0) a.machPart[b] = sum;
}

// One wave per physical boundary face: on a PC-NSCBC face the characteristic
// boundary state (BoundaryState.cpp:996-2241) and its polynomial-correction
// reconstruction, on the other faces (subsonic in/outflow, walls) a pointwise
// boundary state; then the boundary flux (FluxesBoundaryFace,
// FluxesFace.cpp:16-214) and its lift into the element's face slot. The
// branch is uniform over the workgroup (one face each).
template <int K, int M>
__global__ __launch_bounds__(64) void nscbc_kernel(BndArgs<K, M> a)
{
  constexpr int ND = K * K * K, N2 = M * M;
  __shared__ BndSmem<K, M> sm;
  const int b = blockIdx.x, tid = threadIdx.x;
  const int e = a.bface[2 * b], f = a.bface[2 * b + 1];
  const int dir = f / 2, side = f % 2;
  const double fn = side ? 1.0 : -1.0;
  const int bc = a.nb.bc[f];
  bnd_load_ops(sm, a, side);
  bnd_trace(sm, a.U + (size_t) e * 5 * ND, dir, side, true);
  const double *g = a.bgeo + (size_t) b * 13 * N2;
  // parametric slot (0 r, 1 s, 2 t) of the normal, fast and slow face axes
  const int pN = dir, pF = dir == 0 ? 1 : 0, pS = dir == 2 ? 1 : 2;
  const bool nscbc = bc <= kBCInflowTotalNSCBC;
  if (nscbc) {
    // boundary-averaged Mach number (fixed summation order)
    double mach = 0.0;
    for (int i = a.nb.begin[f] + tid; i < a.nb.begin[f + 1]; i += 64) mach += a.machPart[i];
    mach = wave_sum(mach) * a.nb.invArea[f];
    const double dell = fn * a.nb.dell0;
    const int iPhi = fn > 0.0 ? 0 : 4, iPsi = iPhi == 0 ? 4 : 0;

    for (int m = tid; m < N2; m += 64) {
      // entropy -> primitive state and derivatives (BoundaryState.cpp:1002-1108)
      double V[5], Q[5], dQ[3][5];
  #pragma unroll
      for (int v = 0; v < 5; ++v) V[v] = sm.S[0][v * N2 + m];
      const double V4Inv = 1.0 / V[4];
      const double u = -V4Inv * V[1], v = -V4Inv * V[2], w = -V4Inv * V[3];
      const double ek = 0.5 * (u * u + v * v + w * w);
      const double sEnt = kGam - kGm1 * (V[0] - V[4] * ek);
      const double a2 = -kGam * V4Inv;
      const double rho = pow_m2p5(-V[4] * exp(sEnt));
      const double p = a2 * rho / kGam;
      const double povrho = -V4Inv;
      const double E = povrho / kGm1 + ek, H = E + povrho;
  #pragma unroll
      for (int c = 0; c < 3; ++c) {
        const double *dV = sm.S[1 + c];
        const double d0 = dV[m], d1 = dV[N2 + m], d2 = dV[2 * N2 + m], d3 = dV[3 * N2 + m], d4 = dV[4 * N2 + m];
        dQ[c][0] = rho * d0 + rho * u * d1 + rho * v * d2 + rho * w * d3 + rho * E * d4;
        dQ[c][1] = povrho * d1 + povrho * u * d4;
        dQ[c][2] = povrho * d2 + povrho * v * d4;
        dQ[c][3] = povrho * d3 + povrho * w * d4;
        dQ[c][4] = p * d0 + p * u * d1 + p * v * d2 + p * w * d3 + p * H * d4;
      }
      Q[0] = rho; Q[1] = u; Q[2] = v; Q[3] = w; Q[4] = p;
      // normal wave amplitudes with the incoming ones replaced
      const double rr[3] = {g[(4 + 3 * pN) * N2 + m], g[(5 + 3 * pN) * N2 + m], g[(6 + 3 * pN) * N2 + m]};
      const double magrr = sqrt(rr[0] * rr[0] + rr[1] * rr[1] + rr[2] * rr[2]);
      const double ovmagrr = 1.0 / magrr;
      const double rn[3] = {rr[0] * ovmagrr, rr[1] * ovmagrr, rr[2] * ovmagrr};
      const double drhodr = dQ[0][0], dudr = dQ[0][1], dvdr = dQ[0][2], dwdr = dQ[0][3], dpdr = dQ[0][4];
      const double ovrho = 1.0 / rho, c2 = kGam * p * ovrho, c = sqrt(c2);
      const double ova2 = 1.0 / c2, rhoa = rho * c, ovrhoa = 1.0 / rhoa, rhova = rho / c;
      const double dsdr = drhodr - ova2 * dpdr;
      const double dw1dr = rn[1] * dwdr - rn[2] * dvdr;
      const double dw2dr = rn[2] * dudr - rn[0] * dwdr;
      const double dw3dr = rn[0] * dvdr - rn[1] * dudr;
      const double durdr = rn[0] * dudr + rn[1] * dvdr + rn[2] * dwdr;
      double LL[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
      double lmb1, lmb3;
      if (bc == kBCOutflowNSCBC) {
        // BoundaryState.cpp:1318-1519
        const double sign = iPhi == 0 ? -1.0 : 1.0;
        const double kk = 0.5 * a.nb.outlet_sigma * (1.0 - mach * mach) / a.nb.outlet_len;
        const double bl = a.nb.outlet_beta_l < 0.0 ? mach : a.nb.outlet_beta_l;
        const double bt = a.nb.outlet_beta_t < 0.0 ? mach : a.nb.outlet_beta_t;
        const double cbl = a.nb.outlet_one_minus_beta ? fmin(1.0, 1.0 - bl) : fmin(1.0, bl);
        const double cbt = a.nb.outlet_one_minus_beta ? fmin(1.0, 1.0 - bt) : fmin(1.0, bt);
        const double ss[3] = {g[(4 + 3 * pF) * N2 + m], g[(5 + 3 * pF) * N2 + m], g[(6 + 3 * pF) * N2 + m]};
        const double tt[3] = {g[(4 + 3 * pS) * N2 + m], g[(5 + 3 * pS) * N2 + m], g[(6 + 3 * pS) * N2 + m]};
        const double duds = dQ[1][1], dudt = dQ[2][1], dvds = dQ[1][2], dvdt = dQ[2][2];
        const double dwds = dQ[1][3], dwdt = dQ[2][3], dpds = dQ[1][4], dpdt = dQ[2][4];
        const double ur = u * rr[0] + v * rr[1] + w * rr[2];
        const double us = u * ss[0] + v * ss[1] + w * ss[2];
        const double ut = u * tt[0] + v * tt[1] + w * tt[2];
        const double arr = c * magrr;
        const double ars = c * (rn[0] * ss[0] + rn[1] * ss[1] + rn[2] * ss[2]);
        const double art = c * (rn[0] * tt[0] + rn[1] * tt[1] + rn[2] * tt[2]);
        lmb1 = ur - arr; lmb3 = ur + arr;
        const double durds = rn[0] * duds + rn[1] * dvds + rn[2] * dwds;
        const double durdt = rn[0] * dudt + rn[1] * dvdt + rn[2] * dwdt;
        LL[0] = 0.5 * lmb1 * (ovrhoa * dpdr - durdr);
        LL[1] = rn[0] * dsdr - dw1dr;
        LL[2] = rn[1] * dsdr - dw2dr;
        LL[3] = rn[2] * dsdr - dw3dr;
        LL[4] = 0.5 * lmb3 * (ovrhoa * dpdr + durdr);
        const double LT = 0.5 * ovrhoa * (us + sign * ars) * dpds + sign * 0.5 * us * durds
                        + 0.5 * ovrhoa * (ut + sign * art) * dpdt + sign * 0.5 * ut * durdt;
        const double TT = -0.5 * c * (ss[0] * duds + ss[1] * dvds + ss[2] * dwds
                                    + tt[0] * dudt + tt[1] * dvdt + tt[2] * dwdt);
        LL[iPhi] = kk * ovrho * magrr * (p - a.nb.data[f][0]) - cbl * LT + cbt * TT;
      } else if (bc == kBCInflowStaticNSCBC) {
        // BoundaryState.cpp:1596-1783
        const double sign = fn > 0.0 ? -1.0 : 1.0;
        const double ka = sign * 0.5 * a.nb.inlet_sigma * (1.0 - mach * mach) / a.nb.inlet_len;
        const double ku = a.nb.inlet_sigma / a.nb.inlet_len;
        const double ur = u * rr[0] + v * rr[1] + w * rr[2];
        const double arr = c * magrr;
        lmb1 = ur - arr; lmb3 = ur + arr;
        if (a.nb.inlet_incoming) {
          LL[0] = 0.5 * lmb1 * (ovrhoa * dpdr - durdr);
          LL[4] = 0.5 * lmb3 * (ovrhoa * dpdr + durdr);
        }
        const double drho = rho - a.nb.data[f][0], du = u - a.nb.data[f][1], dv = v - a.nb.data[f][2],
                     dw = w - a.nb.data[f][3];
        const double kaol = ku * c * (1.0 / ur);
        LL[iPhi] = ka * c * (rr[0] * du + rr[1] * dv + rr[2] * dw);
        LL[1] = kaol * (rr[2] * dv - rr[1] * dw + rr[0] * drho);
        LL[2] = kaol * (rr[0] * dw - rr[2] * du + rr[1] * drho);
        LL[3] = kaol * (rr[1] * du - rr[0] * dv + rr[2] * drho);
      } else {
        // total-condition inflow, BoundaryState.cpp:1866-2184
        const double ovrg = 1.0 / 287.058, ovcp = kGm1 / (kGam * 287.058);
        const double kk = -a.nb.inlet_sigma / a.nb.inlet_len;
        const double ova = 1.0 / c, T = p * ovrho * ovrg;
        const double ur = u * rn[0] + v * rn[1] + w * rn[2];
        const double w1 = rn[1] * w - rn[2] * v, w2 = rn[2] * u - rn[0] * w, w3 = rn[0] * v - rn[1] * u;
        lmb1 = ur - c; lmb3 = ur + c;
        const double ovlmb2 = 1.0 / ur;
        if (a.nb.inlet_incoming) {
          LL[0] = 0.5 * lmb1 * (ovrhoa * dpdr - durdr);
          LL[4] = 0.5 * lmb3 * (ovrhoa * dpdr + durdr);
        }
        const double magU2 = 2.0 * ek, magU = sqrt(magU2);
        const double ff = 1.0 + 0.5 * kGm1 * (magU2 / c2);
        const double Pt = p * pow(ff, kGam / kGm1), Tt = T * ff;
        const double Zt = v / magU, Zp = w / magU;
        const double ka = kk * c;
        const double X1 = ka * (Pt - a.nb.data[f][0]), X2 = ka * (Tt - a.nb.data[f][1]);
        const double X3 = ka * (Zt - a.nb.data[f][3]), X4 = ka * (Zp - a.nb.data[f][4]);
        const double ovrPtovTt = ovrg * (Pt / Tt), Ttovrho = Tt * ovrho;
        const double govaPt = kGam * ova * Pt, ekovrho = ek * ovrho, gm1ekova = kGm1 * ek * ova;
        const double gm1ovaTt = kGm1 * ova * Tt, ovmagU2 = 1.0 / magU2;
        const double gp = gm1ekova + ur, gmn = gm1ekova - ur;
        const double f1 = w1 - rn[0] * ekovrho, f2 = w2 - rn[1] * ekovrho, f3 = w3 - rn[2] * ekovrho;
        const double A1 = gp * ovrPtovTt - govaPt, A2 = f1 * ovrPtovTt, A3 = f2 * ovrPtovTt, A4 = f3 * ovrPtovTt;
        const double A5 = gmn * ovrPtovTt - govaPt;
        const double B1 = gp * ovcp - gm1ovaTt, B2 = f1 * ovcp + rn[0] * Ttovrho, B3 = f2 * ovcp + rn[1] * Ttovrho;
        const double B4 = f3 * ovcp + rn[2] * Ttovrho, B5 = gmn * ovcp - gm1ovaTt;
        const double C1 = (rn[1] * magU - ur * Zt) * ovmagU2, C2 = -(rn[2] * magU + w1 * Zt) * ovmagU2;
        const double C3 = (-w2 * Zt) * ovmagU2, C4 = (rn[0] * magU - w3 * Zt) * ovmagU2, C5 = -C1;
        const double D1 = (rn[2] * magU - ur * Zp) * ovmagU2, D2 = (rn[1] * magU - w1 * Zp) * ovmagU2;
        const double D3 = -(rn[0] * magU + w2 * Zp) * ovmagU2, D4 = (-w3 * Zp) * ovmagU2, D5 = -D1;
        const double Y1 = LL[iPsi] * A1, Y2 = LL[iPsi] * B1, Y3 = LL[iPsi] * C1, Y4 = LL[iPsi] * D1;
        // elimination of the 4 x 4 system for the incoming amplitudes
        const double ovA2 = 1.0 / A2;
        const double E0 = ovA2 * X1, E3 = -ovA2 * A3, E4 = -ovA2 * A4, E5 = -ovA2 * A5, P2 = -ovA2 * Y1;
        const double ovdF = 1.0 / (D3 + D2 * E3);
        const double F0 = ovdF * (X4 - D2 * E0), F4 = -ovdF * (D4 + D2 * E4), F5 = -ovdF * (D5 + D2 * E5);
        const double P3 = -ovdF * (Y4 + D2 * P2);
        const double ovdG = 1.0 / (C4 + C2 * E4 + C3 * F4 + C2 * E3 * F4);
        const double G0 = ovdG * (X3 - C2 * E0 - C3 * F0 - C2 * E3 * F0);
        const double G5 = -ovdG * (C5 + C2 * E5 + C3 * F5 + C2 * E3 * F5);
        const double P4 = -ovdG * (Y3 + C2 * P2 + C3 * P3 + C2 * E3 * P3);
        const double ovdH = 1.0 / (B5 + B2 * E5 + B3 * F5 + B4 * G5 + B2 * E3 * F5 + B2 * E4 * G5 + B3 * F4 * G5
                                   + B2 * E3 * F4 * G5);
        const double H0 = ovdH * (X2 - B2 * E0 - B3 * F0 - B4 * G0 - B2 * E3 * F0 - B2 * E4 * G0 - B3 * F4 * G0
                                  - B2 * E3 * F4 * G0);
        const double P5 = -ovdH * (Y2 + B2 * P2 + B3 * P3 + B4 * P4 + B2 * E3 * P3 + B2 * E4 * P4 + B3 * F4 * P4
                                   + B2 * E3 * F4 * P4);
        LL[iPhi] = H0 + P5;
        LL[3] = G0 + G5 * LL[iPhi] + P4;
        LL[2] = F0 + F4 * LL[3] + F5 * LL[iPhi] + P3;
        LL[1] = E0 + E3 * LL[2] + E4 * LL[3] + E5 * LL[iPhi] + P2;
        LL[1] *= ovlmb2; LL[2] *= ovlmb2; LL[3] *= ovlmb2;
      }
      // target normal derivative of the primitive state and the B matrix
      const double ovlmb1 = 1.0 / lmb1, ovlmb3 = 1.0 / lmb3;
      double t[5];
      t[0] = ovlmb1 * rhova * LL[0] + rn[0] * LL[1] + rn[1] * LL[2] + rn[2] * LL[3] + ovlmb3 * rhova * LL[4];
      t[1] = -ovlmb1 * rn[0] * LL[0] - rn[2] * LL[2] + rn[1] * LL[3] + ovlmb3 * rn[0] * LL[4];
      t[2] = -ovlmb1 * rn[1] * LL[0] + rn[2] * LL[1] - rn[0] * LL[3] + ovlmb3 * rn[1] * LL[4];
      t[3] = -ovlmb1 * rn[2] * LL[0] - rn[1] * LL[1] + rn[0] * LL[2] + ovlmb3 * rn[2] * LL[4];
      t[4] = ovlmb1 * rhoa * LL[0] + ovlmb3 * rhoa * LL[4];
  #pragma unroll
      for (int q = 0; q < 5; ++q) {
        sm.Q[1][q * N2 + m] = t[q];
        sm.X[0][q * N2 + m] = t[q] - (dQ[0][q] - Q[q] * dell);
      }
    }
    __syncthreads();
    // polynomial correction Q_b = (P1 x P1) B / dell (sign of the face):
    // fast axis, then slow axis
    for (int idx = tid; idx < 5 * N2; idx += 64) {
      const int row = idx / M, j = idx - row * M;
      double s = 0.0;
  #pragma unroll
      for (int q = 0; q < M; ++q) s += sm.P1[j * M + q] * sm.X[0][row * M + q];
      sm.X[1][idx] = s;
    }
    __syncthreads();
    const double scale = 1.0 / dell;
    for (int idx = tid; idx < 5 * N2; idx += 64) {
      const int v = idx / N2, m = idx - v * N2, i = m / M, j = m - i * M;
      double s = 0.0;
  #pragma unroll
      for (int q = 0; q < M; ++q) s += sm.P1[i * M + q] * sm.X[1][(v * M + q) * M + j];
      sm.Q[0][idx] = s * scale;
    }
    __syncthreads();
    // primitive -> entropy variables for the state and its normal derivative
    for (int m = tid; m < N2; m += 64) {
      const double rho = sm.Q[0][m], u = sm.Q[0][N2 + m], v = sm.Q[0][2 * N2 + m], w = sm.Q[0][3 * N2 + m],
                   p = sm.Q[0][4 * N2 + m];
      const double dr = sm.Q[1][m], du = sm.Q[1][N2 + m], dv = sm.Q[1][2 * N2 + m], dw = sm.Q[1][3 * N2 + m],
                   dp = sm.Q[1][4 * N2 + m];
      const double sEnt = log(p / pow(rho, kGam));
      const double ovrho = 1.0 / rho, ovp = 1.0 / p, rovp = rho * ovp;
      const double ek = 0.5 * (u * u + v * v + w * w);
      const double W1 = rovp * u, W2 = rovp * v, W3 = rovp * w;
      sm.Q[0][m] = -rovp * ek + (kGam - sEnt) / kGm1;
      sm.Q[0][N2 + m] = W1; sm.Q[0][2 * N2 + m] = W2; sm.Q[0][3 * N2 + m] = W3; sm.Q[0][4 * N2 + m] = -rovp;
      const double q0 = kGam / kGm1 * ovrho - ovp * ek, q4 = ovp * (rovp * ek - 1.0 / kGm1);
      sm.Q[1][m] = q0 * dr - W1 * du - W2 * dv - W3 * dw + q4 * dp;
      sm.Q[1][N2 + m] = u * ovp * dr + rovp * du - ovp * W1 * dp;
      sm.Q[1][2 * N2 + m] = v * ovp * dr + rovp * dv - ovp * W2 * dp;
      sm.Q[1][3 * N2 + m] = w * ovp * dr + rovp * dw - ovp * W3 * dp;
      sm.Q[1][4 * N2 + m] = -ovp * dr + ovp * rovp * dp;
    }
    __syncthreads();
    // transverse derivatives of the reconstructed state on the Gauss points;
    // the reference stores the slow-axis derivative in the fast-axis slot
    // (BoundaryState.cpp:1261-1287)
    double *dSlow = a.bug ? sm.Q[2] : sm.Q[3];
    double *dFast = a.bug ? sm.Q[3] : sm.Q[2];
    for (int idx = tid; idx < 5 * N2; idx += 64) {
      const int v = idx / N2, m = idx - v * N2, i = m / M, j = m - i * M;
      const double *q0 = sm.Q[0] + v * N2;
      double ds = 0.0, dt = 0.0;
  #pragma unroll
      for (int k = 0; k < M; ++k) { ds += sm.dL[i * M + k] * q0[k * M + j]; dt += sm.dL[j * M + k] * q0[i * M + k]; }
      dSlow[idx] = ds; dFast[idx] = dt;
    }
    __syncthreads();
  } else {
    // pointwise boundary state (BoundaryState.cpp:40-118, :156-372,
    // :399-604, :636-792, entropy variables); d: the face's non-dimensional
    // prescribed data
    const double *d = a.nb.data[f];
    for (int m = tid; m < N2; m += 64) {
      double V[5], R[5];
#pragma unroll
      for (int v = 0; v < 5; ++v) V[v] = sm.S[0][v * N2 + m];
      const double n0 = g[m], n1 = g[N2 + m], n2 = g[2 * N2 + m];
      if (bc == kBCWallInviscid) {
        const double Vn = V[1] * n0 + V[2] * n1 + V[3] * n2;
        R[0] = V[0] - 0.5 * Vn * Vn / V[4];
        R[1] = V[1] - Vn * n0; R[2] = V[2] - Vn * n1; R[3] = V[3] - Vn * n2;
        R[4] = V[4];
      } else if (bc == kBCWallHeatFlux || bc == kBCWallIsothermal) {
        R[0] = V[0] - 0.5 * (V[1] * V[1] + V[2] * V[2] + V[3] * V[3]) / V[4];
        R[1] = 0.0; R[2] = 0.0; R[3] = 0.0;
        R[4] = bc == kBCWallHeatFlux ? V[4] : d[1] / d[0];   // d[1] = -pRef / (R rhoRef)
      } else if (bc == kBCInflowSubsonic) {
        const double V4Inv = 1.0 / V[4];
        const double u = -V4Inv * V[1], v = -V4Inv * V[2], w = -V4Inv * V[3];
        const double a2 = -kGam * V4Inv, c = sqrt(a2);
        const double nx = fn * n0, ny = fn * n1, nz = fn * n2;
        const double un = u * nx + v * ny + w * nz;
        double riemann = 2.0 * c * kInvGm1 + un;
        const double H = a2 * kInvGm1 + 0.5 * (u * u + v * v + w * w);
        riemann *= sqrt(d[1] / H);
        const double alpha = nx * d[2] + ny * d[3] + nz * d[4];
        const double aa = 1.0 + 0.5 * kGm1 * alpha * alpha;
        const double bb = -kGm1 * alpha * riemann;
        const double cc = 0.5 * kGm1 * riemann * riemann - 2.0 * d[1];
        double dd = bb * bb - 4.0 * aa * cc; dd = sqrt(fmax(0.0, dd));
        double qR = (-bb + dd) / (2.0 * aa); qR = fmax(0.0, qR);
        double qR2 = qR * qR;
        double aR2 = kGm1 * (d[1] - 0.5 * qR2);
        double MR2 = qR2 / aR2; MR2 = fmin(1.0, MR2);
        const double tmp = 1.0 / (1.0 + 0.5 * kGm1 * MR2);
        aR2 = kGm1 * d[1] * tmp;
        qR2 = MR2 * aR2;
        qR = sqrt(qR2);
        const double pR = d[0] * pow(tmp, kGam * kInvGm1);
        const double rhoR = kGam * pR / aR2;
        const double pRInv = 1.0 / pR;
        const double sR = log(pR / pow(rhoR, kGam));
        R[0] = (kGam - sR) * kInvGm1 - 0.5 * pRInv * rhoR * qR2;
        R[1] = rhoR * pRInv * qR * d[2]; R[2] = rhoR * pRInv * qR * d[3]; R[3] = rhoR * pRInv * qR * d[4];
        R[4] = -rhoR * pRInv;
      } else if (bc == kBCInflowSupersonic || bc == kBCFarfield) {   // the prescribed / free-stream state
#pragma unroll
        for (int v = 0; v < 5; ++v) R[v] = d[v];
      } else if (bc == kBCOutflowSupersonic) {  // the interior state
#pragma unroll
        for (int v = 0; v < 5; ++v) R[v] = V[v];
      } else {   // subsonic outflow
        const double V4Inv = 1.0 / V[4];
        const double u = -V4Inv * V[1], v = -V4Inv * V[2], w = -V4Inv * V[3];
        const double eKin = 0.5 * (u * u + v * v + w * w);
        const double sEnt = kGam - kGm1 * (V[0] - V[4] * eKin);
        const double c = sqrt(-kGam * V4Inv);
        const double nx = fn * n0, ny = fn * n1, nz = fn * n2;
        const double un = u * nx + v * ny + w * nz;
        const double riemann = 2.0 * c * kInvGm1 + un;
        const double pR = d[0];
        const double rhoR = pow(pR / exp(sEnt), 1.0 / kGam);
        const double aR = sqrt(kGam * pR / rhoR);
        const double unR = riemann - 2.0 * aR * kInvGm1;
        const double uR = u + (unR - un) * nx, vR = v + (unR - un) * ny, wR = w + (unR - un) * nz;
        const double pRInv = 1.0 / pR;
        R[0] = (kGam - sEnt) * kInvGm1 - 0.5 * pRInv * rhoR * (uR * uR + vR * vR + wR * wR);
        R[1] = rhoR * pRInv * uR; R[2] = rhoR * pRInv * vR; R[3] = rhoR * pRInv * wR;
        R[4] = -rhoR * pRInv;
      }
#pragma unroll
      for (int v = 0; v < 5; ++v) sm.Q[0][v * N2 + m] = R[v];
    }
    __syncthreads();
  }
  // fluxes with the face normal oriented out of the element
  const double lenScale = a.econst[(size_t) e * kElemConst + dir];
  const double tau = a.tauFact / lenScale;
  // the element's own LES length scale (ElementClass.cpp:2788-2790)
  const double lenLES = a.econst[(size_t) e * kElemConst + kElemLenLES];
  double evMax = 0.0;
  const bool wall = bc >= kBCWallHeatFlux && bc <= kBCWallInviscid;
  const bool forces = a.monitor && wall;
  double frc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int m = tid; m < N2; m += 64) {
    double VL[5], VR[5], gx[5], gy[5], gz[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      VL[v] = sm.S[0][v * N2 + m];
      VR[v] = sm.Q[0][v * N2 + m];
      const int pt = a.bug ? (v < N2 ? v : 0) : m;
      double c[9];
#pragma unroll
      for (int q = 0; q < 9; ++q) c[q] = g[(4 + q) * N2 + pt];
      double pl[3], pr[3];
      pl[pN] = sm.S[1][v * N2 + m]; pl[pF] = sm.S[2][v * N2 + m]; pl[pS] = sm.S[3][v * N2 + m];
      const double xL = pl[0] * c[0] + pl[1] * c[3] + pl[2] * c[6];
      const double yL = pl[0] * c[1] + pl[1] * c[4] + pl[2] * c[7];
      const double zL = pl[0] * c[2] + pl[1] * c[5] + pl[2] * c[8];
      if (nscbc) {
        // the interior and the reconstructed gradients averaged
        pr[pN] = sm.Q[1][v * N2 + m]; pr[pF] = sm.Q[2][v * N2 + m]; pr[pS] = sm.Q[3][v * N2 + m];
        const double xR = pr[0] * c[0] + pr[1] * c[3] + pr[2] * c[6];
        const double yR = pr[0] * c[1] + pr[1] * c[4] + pr[2] * c[7];
        const double zR = pr[0] * c[2] + pr[1] * c[5] + pr[2] * c[8];
        gx[v] = 0.5 * (xL + xR); gy[v] = 0.5 * (yL + yR); gz[v] = 0.5 * (zL + zR);
      } else {
        // the interior gradient (FluxesFace.cpp:60-80)
        gx[v] = xL; gy[v] = yL; gz[v] = zL;
      }
    }
    const double nn[3] = {fn * g[m], fn * g[N2 + m], fn * g[2 * N2 + m]};
    const double wA = sm.W2[m] * g[3 * N2 + m];
    double fl[5], FX[5], FY[5], FZ[5];
    double aux[7], ev = 0.0;
    if (a.sgs) {
      ev = face_eddy(a.sgs, lenLES, VL, VR, gx, gy, gz);
      evMax = fmax(evMax, ev);
    }
    point_flux<true>(VL, VR, gx, gy, gz, nn, wA, tau, a.mu, a.thetaHalf, fl, FX, FY, FZ, wall ? 0.0 : 1.0,
                     bc == kBCWallHeatFlux, a.nb.data[f][0], forces ? aux : nullptr, a.sgs != kSgsNone, ev);
    if (forces) {
      // wall forces (FluxesFace.cpp:999-1013, :1590-1612): the averaged
      // pressure and the viscous stress on the normal into the boundary
      const double abv = 0.5 * fn * sm.W2[m] * g[3 * N2 + m] * aux[0];
      frc[0] += abv * g[m]; frc[1] += abv * g[N2 + m]; frc[2] += abv * g[2 * N2 + m];
      const double area = fn * sm.W2[m] * g[3 * N2 + m];
      const double ax = area * g[m], ay = area * g[N2 + m], az = area * g[2 * N2 + m];
      frc[3] -= aux[1] * ax + aux[4] * ay + aux[5] * az;
      frc[4] -= aux[4] * ax + aux[2] * ay + aux[6] * az;
      frc[5] -= aux[5] * ax + aux[6] * ay + aux[3] * az;
    }
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      const int pt = a.bug ? (v < N2 ? v : 0) : m;
      double c[9];
#pragma unroll
      for (int q = 0; q < 9; ++q) c[q] = g[(4 + q) * N2 + pt];
      double par[3];
      par[0] = FX[v] * c[0] + FY[v] * c[1] + FZ[v] * c[2];
      par[1] = FX[v] * c[3] + FY[v] * c[4] + FZ[v] * c[5];
      par[2] = FX[v] * c[6] + FY[v] * c[7] + FZ[v] * c[8];
      sm.X[0][v * N2 + m] = fl[v];
      sm.Sym[0][v * N2 + m] = par[pN];
      sm.Sym[1][v * N2 + m] = par[pF];
      sm.Sym[2][v * N2 + m] = par[pS];
    }
  }
  if (a.eddyPart) {
    evMax = wave_max(evMax);
    if (tid == 0) a.eddyPart[4 * a.nElem + b] = evMax;
  }
  if (a.monitor) {
    // this face's force sums (zero off the walls), summed in a fixed order
    // over the faces by force_sum_kernel
#pragma unroll
    for (int i = 0; i < 6; ++i) frc[i] = wave_sum(frc[i]);
    if (tid == 0)
#pragma unroll
      for (int i = 0; i < 6; ++i) a.forcePart[(size_t) b * 6 + i] = frc[i];
  }
  __syncthreads();
  // lift: fast axis, slow axis, then the face vectors along the normal
  double *T0 = sm.Q[0], *T1 = sm.Q[1], *T2 = sm.Q[2];
  for (int pc = tid; pc < 5 * M; pc += 64) {
    double xf[M], xF[M], xS[M], xN[M];
#pragma unroll
    for (int q = 0; q < M; ++q) {
      xf[q] = sm.X[0][pc * M + q]; xF[q] = sm.Sym[1][pc * M + q]; xS[q] = sm.Sym[2][pc * M + q]; xN[q] = sm.Sym[0][pc * M + q];
    }
    for (int n = 0; n < K; ++n) {
      double p1 = 0.0, p2 = 0.0, p3 = 0.0;
#pragma unroll
      for (int q = 0; q < M; ++q) {
        p1 += sm.A[q * K + n] * xf[q]; p1 += sm.D[q * K + n] * xF[q];
        p2 += sm.A[q * K + n] * xS[q];
        p3 += sm.A[q * K + n] * xN[q];
      }
      T0[pc * K + n] = p1; T1[pc * K + n] = p2; T2[pc * K + n] = p3;
    }
  }
  __syncthreads();
  // slow axis, straight into the face slot: value and normal-derivative
  // coefficient planes [0|1][v][slow][fast]
  double *out = a.faceRes + ((size_t) e * 6 + f) * 10 * K * K;
  for (int pc = tid; pc < 5 * K; pc += 64) {
    const int v = pc / K, fi = pc - v * K;
    double x1[M], x2[M], x3[M];
#pragma unroll
    for (int q = 0; q < M; ++q) {
      const int j = (v * M + q) * K + fi;
      x1[q] = T0[j]; x2[q] = T1[j]; x3[q] = T2[j];
    }
    for (int n = 0; n < K; ++n) {
      double g1 = 0.0, g2 = 0.0;
#pragma unroll
      for (int q = 0; q < M; ++q) { g1 += sm.A[q * K + n] * x1[q] + sm.D[q * K + n] * x2[q]; g2 += sm.A[q * K + n] * x3[q]; }
      out[(v * K + n) * K + fi] = g1; out[5 * K * K + (v * K + n) * K + fi] = g2;
    }
  }
}

// ======================================================= time-step kernel
struct DtArgs {
  CaseDev c;
  TableOffsets to;
  int P;
};

struct DtBatchArgs {
  const CaseDev *cases;
  const WorkItem *work;
  TableOffsets to;
  int P;
};

// In-place contraction of the 5 K^2 lines of sU along the axis of stride ST
// with the K x K matrix L, by one wave: every lane reads all its lines
// before any lane writes.
template <int K, int ST>
__device__ __forceinline__ void lines_inplace(double *sU, const double *L, int lane)
{
  constexpr int NL = 5 * K * K, PP = (NL + 63) / 64;
  double o[PP][K];
#pragma unroll
  for (int t = 0; t < PP; ++t) {
    const int p = lane + 64 * t;
    if (p < NL) {
      const int b = ST == 1 ? p * K : ST == K ? (p / K) * K * K + p % K : (p / (K * K)) * K * K * K + p % (K * K);
      double x[K];
#pragma unroll
      for (int q = 0; q < K; ++q) x[q] = sU[b + q * ST];
#pragma unroll
      for (int n = 0; n < K; ++n) {
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < K; ++q) s += L[n * K + q] * x[q];
        o[t][n] = s;
      }
    }
  }
#pragma unroll
  for (int t = 0; t < PP; ++t) {
    const int p = lane + 64 * t;
    if (p < NL) {
      const int b = ST == 1 ? p * K : ST == K ? (p / K) * K * K + p % K : (p / (K * K)) * K * K * K + p % (K * K);
#pragma unroll
      for (int n = 0; n < K; ++n) sU[b + n * ST] = o[t][n];
    }
  }
}

// Nodal entropy variables of element e and (GRAD) their parametric
// derivatives along r, s and t, in LDS: sV[q][v][ND], q = 0 the values, 1..3
// the derivatives (the modal basis or its derivative along each axis); one
// wave. sL, sDL: the K x K tables phi_q(r_p), phi'_q(r_p) at the nodal DOFs.
template <int K, bool GRAD>
__device__ __forceinline__ void nodal_state(const double *Ue, const double *sL, const double *sDL, double *sV,
                                            int lane)
{
  constexpr int ND = K * K * K, NQ = GRAD ? 4 : 1;
  // the modal state in chunks of up to 16 loads per lane, each chunk issued
  // before its first load is waited for
  constexpr int NL = 5 * ND, NIT = (NL + 63) / 64, CH = NIT < 16 ? NIT : 16;
  for (int c0 = 0; c0 < NIT; c0 += CH) {
    double u[CH];
#pragma unroll
    for (int t = 0; t < CH; ++t) {
      const int i = lane + 64 * (c0 + t);
      u[t] = Ue[i < NL ? i : 0];
    }
#pragma unroll
    for (int t = 0; t < CH; ++t) {
      const int i = lane + 64 * (c0 + t);
      if (i < NL)
#pragma unroll
        for (int q = 0; q < NQ; ++q) sV[q * NL + i] = u[t];
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NQ; ++q) lines_inplace<K, 1>(sV + q * 5 * ND, q == 1 ? sDL : sL, lane);
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NQ; ++q) lines_inplace<K, K>(sV + q * 5 * ND, q == 2 ? sDL : sL, lane);
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NQ; ++q) lines_inplace<K, K * K>(sV + q * 5 * ND, q == 3 ? sDL : sL, lane);
  __syncthreads();
}

// Dimensional eddy viscosity at nodal DOF d from nodal_state<K, true>: the
// dimensional density and velocity gradients of the entropy variables with
// the nodal metric vn (ElementClass::ComputeVelocityGradientsNodalDOFs,
// ElementClass.cpp:1649-1790) and the element's LES length scale
// (ElementClass.cpp:1562-1573, :3785-3797).
template <int K>
__device__ __forceinline__ double nodal_eddy(const CaseDev &a, const double *sV, const double *vn, double lenLES,
                                             int d)
{
  constexpr int ND = K * K * K;
  double V[5], rho, u, v, w, p, ek;
#pragma unroll
  for (int l = 0; l < 5; ++l) V[l] = sV[l * ND + d];
  ent2prim(V, rho, u, v, w, p, ek);
  rho *= a.rho_ref; p *= a.p_ref;
  const double V4Inv = -p / (a.u_ref * a.u_ref) / rho;
  const double vel[3] = {u, v, w};
  double gp[3][3];     // [velocity][r | s | t]
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      const double *D = sV + (1 + r) * 5 * ND;
      gp[i][r] = -V4Inv * (D[(1 + i) * ND + d] + vel[i] * D[4 * ND + d]);
    }
  double G[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int x = 0; x < 3; ++x)
      G[i][x] = a.u_ref * (gp[i][0] * vn[(1 + x) * ND + d] + gp[i][1] * vn[(4 + x) * ND + d]
                           + gp[i][2] * vn[(7 + x) * ND + d]);
  return eddy_viscosity(a.sgs, lenLES, rho, G);
}

// Per-element time step (ElementClass::ComputeTimeStep, ElementClass.cpp:
// 1522-1644); SGS: with the eddy viscosity at the nodal DOFs in the viscous
// spectral radius, with the turbulent Prandtl number.
template <int K, bool SGS>
__device__ __forceinline__ void time_step_body(const CaseDev &a, const TableOffsets &to, const int P, const int e)
{
  constexpr int ND = K * K * K;
  __shared__ double sL[K * K], sDL[SGS ? K * K : 1], sU[(SGS ? 20 : 5) * ND];
  const int tid = threadIdx.x;
  for (int i = tid; i < K * K; i += 64) {
    sL[i] = a.tables[to.LD + i];
    if constexpr (SGS) sDL[i] = a.tables[to.DLD + i];
  }
  // modal -> nodal values (and derivatives) along i, j and k, in place
  nodal_state<K, SGS>(a.U + (size_t) e * 5 * ND, sL, sDL, sU, tid);
  const double *sT1 = sU;
  const double *ec = a.econst + (size_t) e * kElemConst;
  const double di = 1.0 / ec[0], dj = 1.0 / ec[1], dk = 1.0 / ec[2];
  const double dl2 = di * di + dj * dj + dk * dk;
  const double rvn = fmax(fmax(1.0, 2.0 + kLamOverMu), kGam / kPrLam) * dl2;
  const double rvnT = fmax(fmax(1.0, 2.0 + kLamOverMu), kGam / kPrTurb) * dl2;
  const double *vn = SGS ? a.vgeoN + (size_t) e * kNodalGeo * ND : nullptr;
  double radInv = 0.0, radVis = 0.0;
  int badA2 = 0;                     // a^2 negative or not a number at a DOF
  for (int d = tid; d < ND; d += 64) {
    double V[5];
#pragma unroll
    for (int l = 0; l < 5; ++l) V[l] = sT1[l * ND + d];
    double rho, u, v, w, p, ek;
    ent2prim(V, rho, u, v, w, p, ek);
    rho *= a.rho_ref; u *= a.u_ref; v *= a.u_ref; w *= a.u_ref; p *= a.p_ref;
    const double ri = 1.0 / rho, a2 = kGam * p * ri;
    const double vI = u * ec[4] + v * ec[5] + w * ec[6];
    const double vJ = u * ec[7] + v * ec[8] + w * ec[9];
    const double vK = u * ec[10] + v * ec[11] + w * ec[12];
    const double c = sqrt(fabs(a2));
    radInv = fmax(radInv, (fabs(vI) + c) * di + (fabs(vJ) + c) * dj + (fabs(vK) + c) * dk);
    if constexpr (SGS) {
      const double ev = nodal_eddy<K>(a, sU, vn, ec[kElemLenLES], d);
      radVis = fmax(radVis, ri * (a.muDim * rvn + ev * rvnT));
    } else {
      radVis = fmax(radVis, ri * a.muDim * rvn);
    }
    badA2 |= !(a2 >= 0.0);
  }
  radInv = wave_max(radInv); radVis = wave_max(radVis);
  const bool anyBad = __any(badA2);
  if (tid == 0) {
    const double f1[] = {2.0, 6.0, 12.0, 20.0, 28.0, 38.0, 50.0, 64.0, 80.0, 98.0};
    const double f2[] = {6.0, 36.0, 150.0, 420.0, 980.0, 1975.0, 3575.0, 7000.0, 14000.0, 28000.0};
    if (anyBad) atomicOr(a.flags, kFlagNegativeA2);
    const double dt = a.cfl * a.u_ref / (f1[P] * radInv + f2[P] * radVis);
    a.dtElem[e] = dt;
  }
}

template <int K, bool SGS>
__global__ __launch_bounds__(64) void time_step_kernel(DtArgs a)
{
  time_step_body<K, SGS>(a.c, a.to, a.P, blockIdx.x);
}

template <int K, bool SGS>
__global__ __launch_bounds__(64) void time_step_batch_kernel(DtBatchArgs b)
{
  const WorkItem w = b.work[blockIdx.x];
  const CaseDev c = b.cases[w.c];
  if (c.flags[kWordActive] == 0) return;
  time_step_body<K, SGS>(c, b.to, b.P, w.e);
}

// Running averages of the dimensional primitive variables and the velocity
// products at the nodal DOFs (ElementClass::UpdateAverageData,
// ElementClass.cpp:3741-3782): avg = factOld avg + factNew q, with q =
// (rho, u, v, w, p, uu, vv, ww, uv, uw, vw) and, with an SGS model, the eddy
// viscosity (:3785-3803); layout [nElem][11 | 12][nDOF].
struct AvgArgs {
  CaseDev c;
  double *avg;
  TableOffsets to;
  double factNew, factOld;
};

template <int K, bool SGS>
__global__ __launch_bounds__(64) void average_kernel(AvgArgs a)
{
  constexpr int ND = K * K * K, NA = SGS ? 12 : 11;
  __shared__ double sL[K * K], sDL[SGS ? K * K : 1], sU[(SGS ? 20 : 5) * ND];
  const CaseDev &c = a.c;
  const int e = blockIdx.x, tid = threadIdx.x;
  for (int i = tid; i < K * K; i += 64) {
    sL[i] = c.tables[a.to.LD + i];
    if constexpr (SGS) sDL[i] = c.tables[a.to.DLD + i];
  }
  nodal_state<K, SGS>(c.U + (size_t) e * 5 * ND, sL, sDL, sU, tid);
  double *av = a.avg + (size_t) e * NA * ND;
  const double *vn = SGS ? c.vgeoN + (size_t) e * kNodalGeo * ND : nullptr;
  for (int d = tid; d < ND; d += 64) {
    double V[5];
#pragma unroll
    for (int l = 0; l < 5; ++l) V[l] = sU[l * ND + d];
    double rho, u, v, w, p, ek;
    ent2prim(V, rho, u, v, w, p, ek);
    rho *= c.rho_ref; u *= c.u_ref; v *= c.u_ref; w *= c.u_ref; p *= c.p_ref;
    const double q[11] = {rho, u, v, w, p, u * u, v * v, w * w, u * v, u * w, v * w};
#pragma unroll
    for (int i = 0; i < 11; ++i) av[i * ND + d] = a.factOld * av[i * ND + d] + a.factNew * q[i];
    if constexpr (SGS) {
      const double ev = nodal_eddy<K>(c, sU, vn, c.econst[(size_t) e * kElemConst + kElemLenLES], d);
      av[11 * ND + d] = a.factOld * av[11 * ND + d] + a.factNew * ev;
    }
  }
}

// Sponge target from the running averages (ElementClass::InitializeDampingState,
// SpongeLayer.cpp:943-997): the averaged nodal primitive variables
// (dimensional, avg[e][0..4]) interpolated to the integration points with the
// Lagrange basis of the nodal DOFs, made non-dimensional and converted to the
// conservative variables, into dampState of the element's sponge slot. One
// wave per element; elements outside the sponge return at once.
struct DampAvgArgs {
  const double *avg, *tables;
  const int *spIdx;
  double *dampState;
  TableOffsets to;
  double u_ref, p_ref, rho_ref;
  int nAvg;                     // averaged quantities per DOF
};

template <int K, int M>
__global__ __launch_bounds__(64) void damping_avg_kernel(DampAvgArgs a)
{
  constexpr int ND = K * K * K, NI = M * M * M;
  __shared__ double sL[M * K], s0[5 * K * K * M], s1[5 * K * M * M], s2[5 * NI];
  const int e = blockIdx.x, tid = threadIdx.x;
  const int sp = a.spIdx[e];
  if (sp < 0) return;
  for (int i = tid; i < M * K; i += 64) sL[i] = a.tables[a.to.Lq + i];
  __syncthreads();
  const double *av = a.avg + (size_t) e * a.nAvg * ND;
  // along i: s0[l][k][j][p]
  for (int o = tid; o < 5 * K * K * M; o += 64) {
    const int p = o % M, r = o / M;          // r = (l k j)
    const double *x = av + (r / (K * K)) * ND + (r % (K * K)) * K;
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < K; ++q) s += sL[p * K + q] * x[q];
    s0[o] = s;
  }
  __syncthreads();
  // along j: s1[l][k][pj][pi]
  for (int o = tid; o < 5 * K * M * M; o += 64) {
    const int pi = o % M, pj = (o / M) % M, lk = o / (M * M);
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < K; ++q) s += sL[pj * K + q] * s0[(lk * K + q) * M + pi];
    s1[o] = s;
  }
  __syncthreads();
  // along k: s2[l][pk][pj][pi]
  for (int o = tid; o < 5 * NI; o += 64) {
    const int pij = o % (M * M), pk = (o / (M * M)) % M, l = o / NI;
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < K; ++q) s += sL[pk * K + q] * s1[((l * K + q) * M * M) + pij];
    s2[o] = s;
  }
  __syncthreads();
  double *ds = a.dampState + (size_t) sp * 5 * NI;
  for (int m = tid; m < NI; m += 64) {
    const double rho = s2[m] / a.rho_ref, u = s2[NI + m] / a.u_ref, v = s2[2 * NI + m] / a.u_ref,
                 w = s2[3 * NI + m] / a.u_ref, pr = s2[4 * NI + m] / a.p_ref;
    ds[m] = rho;
    ds[NI + m] = rho * u;
    ds[2 * NI + m] = rho * v;
    ds[3 * NI + m] = rho * w;
    ds[4 * NI + m] = pr * kInvGm1 + 0.5 * rho * (u * u + v * v + w * w);
  }
}

// Minimum of the per-element time steps into scal[0] (one workgroup; a
// fixed reduction order, no atomics on one address from every element).
__global__ __launch_bounds__(1024) void dt_min_kernel(const double *dtElem, int n, double *scal)
{
  __shared__ double red[16];
  double m = 1e10;
  for (int i = threadIdx.x; i < n; i += 1024) m = fmin(m, dtElem[i]);
  m = -wave_max(-m);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = red[0];
    for (int w = 1; w < 16; ++w) r = fmin(r, red[w]);
    scal[0] = r;
  }
}

// Per case of a batch group (one workgroup each): the minimum of the
// per-element time steps as dt_min_kernel forms it, then the step control of
// SolverClass::ExecuteTimeSynchronizationStep (SolverClass.cpp:2356-2365):
// the step that reaches 0.99 of the synchronization time step is clipped to
// end it exactly, and the next step starts a new one. A case with error flags
// or no RK steps left is deactivated, so the batch's launches skip it.
__global__ __launch_bounds__(1024) void batch_sync_kernel(const CaseDev *cases)
{
  __shared__ double red[16];
  const CaseDev c = cases[blockIdx.x];
  if (c.flags[kWordActive] == 0) return;
  double m = 1e10;
  for (int i = threadIdx.x; i < c.nElem; i += 1024) m = fmin(m, c.dtElem[i]);
  m = -wave_max(-m);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = red[0];
    for (int w = 1; w < 16; ++w) r = fmin(r, red[w]);
    if (c.flags[kWordFlags] != 0 || c.flags[kWordStepsLeft] <= 0) {
      c.flags[kWordActive] = 0;
      return;
    }
    c.scal[0] = r;
    double dt = r;
    const double el = c.scal[kScalElapsed], dtSync = c.scal[kScalDtSync];
    if (el + dt >= 0.99 * dtSync) {
      dt = dtSync - el;
      c.scal[kScalElapsed] = 0.0;
      c.scal[kScalSyncDone] += 1.0;
    } else {
      c.scal[kScalElapsed] = el + dt;
    }
    c.scal[kScalDt] = dt;
    c.flags[kWordStepsLeft] -= 1;
    c.flags[kWordStepsDone] += 1;
  }
}

// First RK stage of a batch: U0 <- U and U <- U - dt R per element, with the
// case's time step (the same values stage0_kernel forms after its buffer swap).
__global__ __launch_bounds__(256) void stage0_batch_kernel(const CaseDev *cases, const WorkItem *work, int nd5)
{
  const WorkItem w = work[blockIdx.x];
  const CaseDev c = cases[w.c];
  if (c.flags[kWordActive] == 0) return;
  const double dt = c.scal[kScalDt];
  const size_t base = (size_t) w.e * nd5;
  for (int i = threadIdx.x; i < nd5; i += 256) {
    const double u = c.U[base + i];
    c.U0[base + i] = u;
    c.U[base + i] = u - dt * c.R[base + i];
  }
}

// Status words of the cases of a batch group, gathered for one readback. A
// case counts as active while the next RK step of the batch would advance it
// (its own word is cleared only by that step's batch_sync_kernel).
__global__ __launch_bounds__(64) void batch_status_kernel(const CaseDev *cases, int n, int *out)
{
  const int i = blockIdx.x * 64 + threadIdx.x;
  if (i >= n) return;
  const CaseDev c = cases[i];
  out[i * kStatusWords + 0] =
      c.flags[kWordActive] != 0 && c.flags[kWordFlags] == 0 && c.flags[kWordStepsLeft] > 0 ? 1 : 0;
  out[i * kStatusWords + 1] = c.flags[kWordStepsDone];
  out[i * kStatusWords + 2] = c.flags[kWordStepsLeft];
  out[i * kStatusWords + 3] = c.flags[kWordFlags];
  out[i * kStatusWords + 4] = (int) c.scal[kScalSyncDone];
}

// Maximum of n non-negative values into *out (one workgroup).
__global__ __launch_bounds__(1024) void reduce_max_kernel(const double *v, int n, double *out)
{
  __shared__ double red[16];
  double m = 0.0;
  for (int i = threadIdx.x; i < n; i += 1024) m = fmax(m, v[i]);
  m = wave_max(m);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = red[0];
    for (int w = 1; w < 16; ++w) r = fmax(r, red[w]);
    *out = r;
  }
}

// Sums of the six wall-force components over the n boundary faces into
// out[0..6), in face order (one wave, each lane a fixed stride of faces, then
// the lanes in order).
__global__ __launch_bounds__(64) void force_sum_kernel(const double *part, int n, double *out)
{
  __shared__ double red[64][6];
  double s[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int b = threadIdx.x; b < n; b += 64)
#pragma unroll
    for (int i = 0; i < 6; ++i) s[i] += part[(size_t) b * 6 + i];
#pragma unroll
  for (int i = 0; i < 6; ++i) red[threadIdx.x][i] = s[i];
  __syncthreads();
  if (threadIdx.x < 6) {
    double t = 0.0;
    for (int l = 0; l < 64; ++l) t += red[l][threadIdx.x];
    out[threadIdx.x] = t;
  }
}

// First RK stage: U <- U0 - dt R, where U0 already holds the state of the
// step (the host swaps the two buffers instead of copying). dt is the
// argument, which the first lane also stores to dtOut for the later stages,
// or (dtIn non-null) is read from device memory where the host put it.
__global__ void stage0_kernel(double *U, const double *U0, const double *R, size_t n, double dtArg,
                              const double *dtIn, double *dtOut)
{
  const double dt = dtIn ? *dtIn : dtArg;
  if (dtOut && blockIdx.x == 0 && threadIdx.x == 0) *dtOut = dt;
  for (size_t i = blockIdx.x * (size_t) blockDim.x + threadIdx.x; i < n; i += (size_t) gridDim.x * blockDim.x)
    U[i] = U0[i] - dt * R[i];
}

// p[0, n) = v (a kernel, so that it can be captured in a graph).
__global__ void fill_kernel(double *p, size_t n, double v)
{
  for (size_t i = blockIdx.x * (size_t) blockDim.x + threadIdx.x; i < n; i += (size_t) gridDim.x * blockDim.x)
    p[i] = v;
}

// The time step of the RK step about to run, into device memory (one lane,
// a vector store).
__global__ void set_dt_kernel(double *p, double v)
{
  if (threadIdx.x == 0) *p = v;
}

template <int K, int M, typename T>
void fill_face_ops(const DeviceCase &dc, FaceOps<K, M, T> &op)
{
  for (int i = 0; i < M * K; ++i) { op.A[i] = T(dc.opA[i]); op.D[i] = T(dc.opD[i]); }
  for (int q = 0; q < K; ++q)
    for (int t = 0; t < 4; ++t) op.F[t][q] = T(dc.opF[t][q]);
  for (int q = 0; q < M; ++q) op.w[q] = dc.opW[q];
  op.c03 = dc.opC0[0] * dc.opC0[0] * dc.opC0[0];
}

template <int K, int M, typename T>
void fill_vol_shape(const DeviceCase &dc, const StageCtl &ctl, VolArgs<K, M, T> &a)
{
  a.ctl = ctl; a.to = dc.to; a.bug = dc.bug_compat;
  for (int i = 0; i < M * K; ++i) { a.op.A[i] = T(dc.opA[i]); a.op.D[i] = T(dc.opD[i]); }
  for (int q = 0; q < K; ++q) {
    a.c0[q] = dc.opC0[q];
    for (int t = 0; t < 4; ++t) a.fvec[t][q] = dc.opF[t][q];
  }
}

template <int K, int M, bool BATCH, typename T, bool SGS>
void launch_face_kernel(bool bug, int n, const FaceArgs<K, M, T> &a, hipStream_t s)
{
  if (bug)
    hipLaunchKernelGGL((face_kernel<K, M, true, BATCH, T, SGS>), dim3(n, 3), dim3(64), 0, s, a);
  else
    hipLaunchKernelGGL((face_kernel<K, M, false, BATCH, T, SGS>), dim3(n, 3), dim3(64), 0, s, a);
}

template <int K, int M, typename T>
hipError_t faces_impl_t(const DeviceCase &dc, hipStream_t s)
{
  FaceArgs<K, M, T> a{};
  a.c = case_dev(dc);
  fill_face_ops(dc, a.op);
  if (dc.sgs) launch_face_kernel<K, M, false, T, true>(dc.bug_compat, dc.nElem, a, s);
  else launch_face_kernel<K, M, false, T, false>(dc.bug_compat, dc.nElem, a, s);
  return hipGetLastError();
}

template <int K, int M>
hipError_t faces_impl(const DeviceCase &dc, hipStream_t s)
{
  return dc.mixed ? faces_impl_t<K, M, float>(dc, s) : faces_impl_t<K, M, double>(dc, s);
}

template <int K, int M, typename T>
hipError_t batch_faces_impl_t(const BatchGroup &g, int w0, int nw, hipStream_t s)
{
  FaceArgs<K, M, T> a{};
  a.cases = g.cases; a.work = g.work; a.w0 = w0;
  fill_face_ops(*g.proto, a.op);
  if (g.proto->sgs) launch_face_kernel<K, M, true, T, true>(g.proto->bug_compat, nw, a, s);
  else launch_face_kernel<K, M, true, T, false>(g.proto->bug_compat, nw, a, s);
  return hipGetLastError();
}

template <int K, int M>
hipError_t batch_faces_impl(const BatchGroup &g, int w0, int nw, hipStream_t s)
{
  return g.proto->mixed ? batch_faces_impl_t<K, M, float>(g, w0, nw, s)
                        : batch_faces_impl_t<K, M, double>(g, w0, nw, s);
}

template <int K, int M>
hipError_t boundary_impl(const DeviceCase &dc, hipStream_t s, bool monitor)
{
  if (dc.nBFace == 0) return hipSuccess;
  BndArgs<K, M> a;
  a.U = dc.U; a.tables = dc.tables; a.econst = dc.econst; a.bgeo = dc.bgeo; a.bface = dc.bface;
  a.machPart = dc.machPart; a.faceRes = dc.faceRes; a.forcePart = dc.forcePart; a.to = dc.to;
  a.bug = dc.bug_compat; a.monitor = monitor && dc.monitor_forces;
  a.mu = dc.mu; a.tauFact = (dc.P + 1.0) * (dc.P + 1.0) * dc.rel_penalty; a.thetaHalf = 0.5 * dc.theta_sym;
  a.nb = dc.nscbc;
  a.sgs = dc.sgs; a.nElem = dc.nElem; a.eddyPart = dc.sgs ? dc.eddyPart : nullptr;
  hipLaunchKernelGGL((mach_kernel<K, M>), dim3(dc.nBFace), dim3(64), 0, s, a);
  hipLaunchKernelGGL((nscbc_kernel<K, M>), dim3(dc.nBFace), dim3(64), 0, s, a);
  if (a.monitor)
    hipLaunchKernelGGL(force_sum_kernel, dim3(1), dim3(64), 0, s, dc.forcePart, dc.nBFace, dc.scal + kScalForce);
  return hipGetLastError();
}

// Volume workgroup sizes instantiated per (K, M), the default first: the
// sizes for which the points fit two passes and the DOFs one thread each.
// The defaults are the autotuner's choices on gfx950 (64 threads at (3, 4)
// and 128 at (4, 5) run 2.4x and 1.6x faster than 256; the recorded table
// is vcp3d_open_amd/tiles_gfx950.json).
#define VCP_TILES(X)                                                            \
  X(3, 4, 64) X(3, 4, 128) X(3, 4, 256) X(4, 5, 128) X(4, 5, 256)              \
  X(4, 6, 256) X(4, 6, 128) X(5, 7, 256) X(5, 7, 512) X(6, 8, 256) X(6, 8, 512) \
  X(7, 9, 512)

template <int K, int M, int NT, typename T>
hipError_t volume_nt_t(const DeviceCase &dc, const StageCtl &ctl, hipStream_t s)
{
  VolArgs<K, M, T> a{};
  a.c = case_dev(dc);
  fill_vol_shape(dc, ctl, a);
  hipLaunchKernelGGL((volume_kernel<K, M, NT, false, T>), dim3(dc.nElem), dim3(NT), 0, s, a);
  return hipGetLastError();
}

template <int K, int M, int NT>
hipError_t volume_nt(const DeviceCase &dc, const StageCtl &ctl, hipStream_t s)
{
  return dc.mixed ? volume_nt_t<K, M, NT, float>(dc, ctl, s) : volume_nt_t<K, M, NT, double>(dc, ctl, s);
}

template <int K, int M, int NT, typename T>
hipError_t batch_volume_nt_t(const BatchGroup &g, const StageCtl &ctl, int w0, int nw, hipStream_t s)
{
  VolArgs<K, M, T> a{};
  a.cases = g.cases; a.work = g.work; a.w0 = w0;
  fill_vol_shape(*g.proto, ctl, a);
  hipLaunchKernelGGL((volume_kernel<K, M, NT, true, T>), dim3(nw), dim3(NT), 0, s, a);
  return hipGetLastError();
}

template <int K, int M, int NT>
hipError_t batch_volume_nt(const BatchGroup &g, const StageCtl &ctl, int w0, int nw, hipStream_t s)
{
  return g.proto->mixed ? batch_volume_nt_t<K, M, NT, float>(g, ctl, w0, nw, s)
                        : batch_volume_nt_t<K, M, NT, double>(g, ctl, w0, nw, s);
}

template <int K>
hipError_t dt_impl(const DeviceCase &dc, hipStream_t s)
{
  DtArgs a;
  a.c = case_dev(dc); a.to = dc.to; a.P = dc.P;
  if (dc.sgs)
    hipLaunchKernelGGL((time_step_kernel<K, true>), dim3(dc.nElem), dim3(64), 0, s, a);
  else
    hipLaunchKernelGGL((time_step_kernel<K, false>), dim3(dc.nElem), dim3(64), 0, s, a);
  hipLaunchKernelGGL(dt_min_kernel, dim3(1), dim3(1024), 0, s, dc.dtElem, dc.nElem, dc.scal);
  return hipGetLastError();
}

template <int K>
hipError_t batch_dt_impl(const BatchGroup &g, hipStream_t s)
{
  DtBatchArgs b;
  b.cases = g.cases; b.work = g.work; b.to = g.proto->to; b.P = g.proto->P;
  if (g.proto->sgs)
    hipLaunchKernelGGL((time_step_batch_kernel<K, true>), dim3(g.nWork), dim3(64), 0, s, b);
  else
    hipLaunchKernelGGL((time_step_batch_kernel<K, false>), dim3(g.nWork), dim3(64), 0, s, b);
  hipLaunchKernelGGL(batch_sync_kernel, dim3(g.nCase), dim3(1024), 0, s, g.cases);
  return hipGetLastError();
}

}  // namespace

CaseDev case_dev(const DeviceCase &dc)
{
  CaseDev c{};
  c.U = dc.U; c.U0 = dc.U0; c.R = dc.R; c.faceRes = dc.faceRes;
  c.tables = dc.tables; c.vgeo = dc.vgeo; c.fgeo = dc.fgeo; c.econst = dc.econst; c.nbr = dc.nbr;
  c.scal = dc.scal; c.flags = dc.flags; c.cgIter = dc.cgIter; c.dtElem = dc.dtElem;
  c.mach2Elem = dc.mach2Elem;
  c.spIdx = dc.nSponge ? dc.spIdx : nullptr; c.spCoef = dc.spCoef; c.dampFn = dc.dampFn;
  c.dampState = dc.dampState; c.srcBuf = dc.srcBuf;
  c.mu = dc.mu; c.tauFact = (dc.P + 1.0) * (dc.P + 1.0) * dc.rel_penalty; c.thetaHalf = 0.5 * dc.theta_sym;
  c.cfl = dc.cfl; c.u_ref = dc.u_ref; c.p_ref = dc.p_ref; c.rho_ref = dc.rho_ref;
  c.muDim = dc.mu * dc.p_ref / dc.u_ref;
  c.nElem = dc.nElem; c.cml = dc.cml;
  c.sgs = dc.sgs; c.vgeoN = dc.vgeoN; c.eddyPart = dc.sgs ? dc.eddyPart : nullptr;
  c.bfBuf = dc.fbodyOn ? dc.bfBuf : nullptr;
  for (int i = 0; i < 3; ++i) c.fbody[i] = dc.fbody[i];
  return c;
}

#define VCP_PAIRS(X) X(3, 4) X(4, 5) X(4, 6) X(5, 7) X(6, 8) X(7, 9)

bool supported_pair(int K, int M)
{
#define X(k, m) if (K == k && M == m) return true;
  VCP_PAIRS(X)
#undef X
  return false;
}

int volume_tiles(int K, int M, int out[4])
{
  int n = 0;
#define X(k, m, nt) if (K == k && M == m && n < 4) out[n++] = nt;
  VCP_TILES(X)
#undef X
  return n;
}

int default_volume_tile(int K, int M)
{
  int t[4];
  return volume_tiles(K, M, t) > 0 ? t[0] : 0;
}

hipError_t launch_faces(const DeviceCase &dc, hipStream_t s)
{
#define X(k, m) if (dc.K == k && dc.M == m) return faces_impl<k, m>(dc, s);
  VCP_PAIRS(X)
#undef X
  return hipErrorInvalidValue;
}

hipError_t launch_boundary(const DeviceCase &dc, hipStream_t s, bool monitor)
{
#define X(k, m) if (dc.K == k && dc.M == m) return boundary_impl<k, m>(dc, s, monitor);
  VCP_PAIRS(X)
#undef X
  return hipErrorInvalidValue;
}

hipError_t launch_volume(const DeviceCase &dc, const StageCtl &ctl, hipStream_t s)
{
  const int nt = dc.nt ? dc.nt : default_volume_tile(dc.K, dc.M);
  const hipError_t e = [&]() -> hipError_t {
#define X(k, m, t) if (dc.K == k && dc.M == m && nt == t) return volume_nt<k, m, t>(dc, ctl, s);
    VCP_TILES(X)
#undef X
    return hipErrorInvalidValue;
  }();
  if (e == hipSuccess && ctl.monitor) {
    hipLaunchKernelGGL(reduce_max_kernel, dim3(1), dim3(1024), 0, s, dc.mach2Elem, dc.nElem, dc.scal + 1);
    // volume, internal-face and boundary-face maxima of the eddy viscosity
    if (dc.sgs)
      hipLaunchKernelGGL(reduce_max_kernel, dim3(1), dim3(1024), 0, s, dc.eddyPart, 4 * dc.nElem + dc.nBFace,
                         dc.scal + kScalEddy);
  }
  return e == hipSuccess ? hipGetLastError() : e;
}

hipError_t launch_time_step(const DeviceCase &dc, hipStream_t s)
{
  switch (dc.K) {
    case 3: return dt_impl<3>(dc, s);
    case 4: return dt_impl<4>(dc, s);
    case 5: return dt_impl<5>(dc, s);
    case 6: return dt_impl<6>(dc, s);
    case 7: return dt_impl<7>(dc, s);
  }
  return hipErrorInvalidValue;
}

template <int K>
hipError_t avg_impl(const DeviceCase &dc, double factNew, hipStream_t s)
{
  AvgArgs a;
  a.c = case_dev(dc); a.avg = dc.avg; a.to = dc.to;
  a.factNew = factNew; a.factOld = 1.0 - factNew;
  if (dc.sgs)
    hipLaunchKernelGGL((average_kernel<K, true>), dim3(dc.nElem), dim3(64), 0, s, a);
  else
    hipLaunchKernelGGL((average_kernel<K, false>), dim3(dc.nElem), dim3(64), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_average(const DeviceCase &dc, double factNew, hipStream_t s)
{
  switch (dc.K) {
    case 3: return avg_impl<3>(dc, factNew, s);
    case 4: return avg_impl<4>(dc, factNew, s);
    case 5: return avg_impl<5>(dc, factNew, s);
    case 6: return avg_impl<6>(dc, factNew, s);
    case 7: return avg_impl<7>(dc, factNew, s);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_damping_from_averages(const DeviceCase &dc, hipStream_t s)
{
  if (!dc.avg || !dc.nSponge) return hipErrorInvalidValue;
  DampAvgArgs a;
  a.avg = dc.avg; a.tables = dc.tables; a.spIdx = dc.spIdx; a.dampState = dc.dampState; a.to = dc.to;
  a.u_ref = dc.u_ref; a.p_ref = dc.p_ref; a.rho_ref = dc.rho_ref; a.nAvg = dc.nAvg;
#define X(k, m)                                                                                    \
  if (dc.K == k && dc.M == m) {                                                                    \
    hipLaunchKernelGGL((damping_avg_kernel<k, m>), dim3(dc.nElem), dim3(64), 0, s, a);             \
    return hipGetLastError();                                                                      \
  }
  VCP_PAIRS(X)
#undef X
  return hipErrorInvalidValue;
}

hipError_t launch_fill(double *p, size_t n, double v, hipStream_t s)
{
  if (n == 0) return hipSuccess;
  const int blocks = (int) ((n + 255) / 256 < 4096 ? (n + 255) / 256 : 4096);
  hipLaunchKernelGGL(fill_kernel, dim3(blocks), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

hipError_t static_scratch(void **ptr, size_t *bytes)
{
  // one address per device, looked up once with that device current
  static std::mutex mu;
  static void *addr[kMaxDevices] = {};
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[dev]) {
    void *a = nullptr;
    e = hipGetSymbolAddress(&a, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    addr[dev] = a;
  }
  *ptr = addr[dev];
  *bytes = sizeof(scratch);
  return hipSuccess;
}

hipError_t launch_set_dt(const DeviceCase &dc, double dt, hipStream_t s)
{
  hipLaunchKernelGGL(set_dt_kernel, dim3(1), dim3(64), 0, s, dc.scal + kScalDt, dt);
  return hipGetLastError();
}

hipError_t launch_stage0(const DeviceCase &dc, const double *dt, hipStream_t s)
{
  const size_t n = (size_t) dc.nElem * 5 * dc.nDOF;
  const int blocks = (int) ((n + 255) / 256 < 65536 ? (n + 255) / 256 : 65536);
  hipLaunchKernelGGL(stage0_kernel, dim3(blocks), dim3(256), 0, s, dc.U, dc.U0, dc.R, n,
                     dt ? *dt : 0.0, dt ? nullptr : dc.scal + kScalDt, dt ? dc.scal + kScalDt : nullptr);
  return hipGetLastError();
}

hipError_t launch_batch_faces(const BatchGroup &g, int w0, int nw, hipStream_t s)
{
  if (nw <= 0) return hipSuccess;
  const DeviceCase &p = *g.proto;
#define X(k, m) if (p.K == k && p.M == m) return batch_faces_impl<k, m>(g, w0, nw, s);
  VCP_PAIRS(X)
#undef X
  return hipErrorInvalidValue;
}

hipError_t launch_batch_volume(const BatchGroup &g, const StageCtl &ctl, int w0, int nw, hipStream_t s)
{
  if (nw <= 0) return hipSuccess;
  const DeviceCase &p = *g.proto;
  const int nt = g.nt ? g.nt : default_volume_tile(p.K, p.M);
#define X(k, m, t) if (p.K == k && p.M == m && nt == t) return batch_volume_nt<k, m, t>(g, ctl, w0, nw, s);
  VCP_TILES(X)
#undef X
  return hipErrorInvalidValue;
}

hipError_t launch_batch_time_step(const BatchGroup &g, hipStream_t s)
{
  if (g.nWork <= 0) return hipSuccess;
  switch (g.proto->K) {
    case 3: return batch_dt_impl<3>(g, s);
    case 4: return batch_dt_impl<4>(g, s);
    case 5: return batch_dt_impl<5>(g, s);
    case 6: return batch_dt_impl<6>(g, s);
    case 7: return batch_dt_impl<7>(g, s);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_batch_status(const BatchGroup &g, int *out, hipStream_t s)
{
  if (g.nCase <= 0) return hipSuccess;
  hipLaunchKernelGGL(batch_status_kernel, dim3((g.nCase + 63) / 64), dim3(64), 0, s, g.cases, g.nCase, out);
  return hipGetLastError();
}

hipError_t launch_batch_stage0(const BatchGroup &g, hipStream_t s)
{
  if (g.nWork <= 0) return hipSuccess;
  hipLaunchKernelGGL(stage0_batch_kernel, dim3(g.nWork), dim3(256), 0, s, g.cases, g.work,
                     5 * g.proto->nDOF);
  return hipGetLastError();
}

}  // namespace vcp
