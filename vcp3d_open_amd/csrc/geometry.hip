// Metric terms of a block's elements on the device, in the layouts the
// residual kernels read (ElementClass::ComputeMetricTerms, ElementClass.cpp:
// 1027-1407; FinalMetric{I,J,K}Face, :1907-2088; ComputeLengthScales,
// :927-1022). One workgroup per element. The arithmetic repeats the host
// set-up (setup.cpp: build_geometry) operation for operation: the same
// sum-factorised contractions in the same order, and this file is built
// without contracting multiply-adds into fused operations, as the host build
// is, so the device arrays equal the host set-up's bit for bit.
#include "geometry.hpp"

#include "kernels.hpp"

namespace vcp {
namespace {

constexpr int kGeoThreads = 256;
constexpr int kMaxKg = 10;   // grid degree <= 9

template <int M>
struct GeoSmem {
  static constexpr int NI = M * M * M, N2 = M * M, KG3 = kMaxKg * kMaxKg * kMaxKg;
  double g[3][KG3];                   // grid DOFs of the element, per coordinate
  double a[2][kMaxKg * kMaxKg * M];   // contraction along x (up to two operators)
  double b[3][kMaxKg * M * M];        // along y
  double d[9][NI];                    // dx_c/dr, dx_c/ds, dx_c/dt at the points
  double J[NI];
  double fc[9][N2];                   // face: dx_c/dr, /ds, /dt at the face points
  double fn[6][4][N2];                // face normals and area factors
};

// out[k][j][i] = sum_q op[i][q] in[k][j][q], over ni x nj rows of nq inputs
__device__ inline void along_x(const double *op, const double *in, double *out, int nk, int nj, int nq, int no)
{
  for (int t = threadIdx.x; t < nk * nj * no; t += kGeoThreads) {
    const int i = t % no, kj = t / no;
    double s = 0.0;
    for (int q = 0; q < nq; ++q) s += op[i * nq + q] * in[kj * nq + q];
    out[kj * no + i] = s;
  }
}

// out[k][j][i] = sum_q op[j][q] in[k][q][i]
__device__ inline void along_y(const double *op, const double *in, double *out, int nk, int nq, int no, int ni)
{
  for (int t = threadIdx.x; t < nk * no * ni; t += kGeoThreads) {
    const int i = t % ni, j = (t / ni) % no, k = t / (ni * no);
    double s = 0.0;
    for (int q = 0; q < nq; ++q) s += op[j * nq + q] * in[(k * nq + q) * ni + i];
    out[(k * no + j) * ni + i] = s;
  }
}

// out[k][j][i] = sum_q op[k][q] in[q][j][i]
__device__ inline void along_z(const double *op, const double *in, double *out, int nq, int no, int nji)
{
  for (int t = threadIdx.x; t < no * nji; t += kGeoThreads) {
    const int ji = t % nji, k = t / nji;
    double s = 0.0;
    for (int q = 0; q < nq; ++q) s += op[k * nq + q] * in[q * nji + ji];
    out[k * nji + ji] = s;
  }
}

__device__ inline void inverse_metric(const double dr[3], const double ds[3], const double dt[3], double &J,
                                      double m[9])
{
  J = dr[0] * (ds[1] * dt[2] - ds[2] * dt[1]) - ds[0] * (dr[1] * dt[2] - dr[2] * dt[1])
    + dt[0] * (dr[1] * ds[2] - dr[2] * ds[1]);
  const double Ji = 1.0 / J;
  m[0] = Ji * (ds[1] * dt[2] - ds[2] * dt[1]);
  m[1] = Ji * (ds[2] * dt[0] - ds[0] * dt[2]);
  m[2] = Ji * (ds[0] * dt[1] - ds[1] * dt[0]);
  m[3] = Ji * (dr[2] * dt[1] - dr[1] * dt[2]);
  m[4] = Ji * (dr[0] * dt[2] - dr[2] * dt[0]);
  m[5] = Ji * (dr[1] * dt[0] - dr[0] * dt[1]);
  m[6] = Ji * (dr[1] * ds[2] - dr[2] * ds[1]);
  m[7] = Ji * (dr[2] * ds[0] - dr[0] * ds[2]);
  m[8] = Ji * (dr[0] * ds[1] - dr[1] * ds[0]);
}

template <int M>
__global__ __launch_bounds__(kGeoThreads) void geometry_kernel(GeoArgs a)
{
  using S = GeoSmem<M>;
  constexpr int NI = S::NI, N2 = S::N2;
  __shared__ S sm;
  const int e = blockIdx.x, tid = threadIdx.x, Kg = a.Kg, nG = Kg * Kg * Kg;
  const double *Lg = a.tab, *dLg = Lg + M * Kg, *lf = dLg + M * Kg, *dlf = lf + 2 * Kg;
  const double *w2 = dlf + 2 * Kg, *w3 = w2 + N2;
  for (int t = tid; t < 3 * nG; t += kGeoThreads) sm.g[t / nG][t % nG] = a.grid[(size_t) e * 3 * nG + t];
  __syncthreads();
  // ---- volume: dx_c/dr = Lz Ly dLx x_c, dx_c/ds = Lz dLy Lx x_c, dx_c/dt = dLz Ly Lx x_c
  for (int c = 0; c < 3; ++c) {
    along_x(Lg, sm.g[c], sm.a[0], Kg, Kg, Kg, M);
    along_x(dLg, sm.g[c], sm.a[1], Kg, Kg, Kg, M);
    __syncthreads();
    along_y(Lg, sm.a[0], sm.b[0], Kg, Kg, M, M);
    along_y(Lg, sm.a[1], sm.b[1], Kg, Kg, M, M);
    along_y(dLg, sm.a[0], sm.b[2], Kg, Kg, M, M);
    __syncthreads();
    along_z(Lg, sm.b[1], sm.d[0 + c], Kg, M, N2);
    along_z(Lg, sm.b[2], sm.d[3 + c], Kg, M, N2);
    along_z(dLg, sm.b[0], sm.d[6 + c], Kg, M, N2);
    __syncthreads();
  }
  double *vg = a.vgeo + (size_t) e * kVolGeo * NI;
  for (int m = tid; m < NI; m += kGeoThreads) {
    const double dr[3] = {sm.d[0][m], sm.d[1][m], sm.d[2][m]};
    const double ds[3] = {sm.d[3][m], sm.d[4][m], sm.d[5][m]};
    const double dt[3] = {sm.d[6][m], sm.d[7][m], sm.d[8][m]};
    double J, mi[9];
    inverse_metric(dr, ds, dt, J, mi);
    if (!(J > 0.0)) atomicOr(a.flag, 1);
    sm.J[m] = J;
    vg[m] = J;
    for (int q = 0; q < 9; ++q) vg[(1 + q) * NI + m] = mi[q];
  }
  __syncthreads();
  // ---- faces: the grid derivatives at the face points, with the face
  // vector phi(+-1) or phi'(+-1) in the normal slot
  for (int f = 0; f < 6; ++f) {
    const int dir = f / 2, side = f % 2;
    const double *fv = lf + side * Kg, *dfv = dlf + side * Kg;
    for (int op = 1; op < 4; ++op) {            // d/dr, d/ds, d/dt
      const double *Ox = op == 1 ? dLg : Lg, *Oy = op == 2 ? dLg : Lg, *Oz = op == 3 ? dLg : Lg;
      int ox = M, oy = M, oz = M;
      if (dir == 0) { ox = 1; Ox = op == 1 ? dfv : fv; }
      if (dir == 1) { oy = 1; Oy = op == 2 ? dfv : fv; }
      if (dir == 2) { oz = 1; Oz = op == 3 ? dfv : fv; }
      for (int c = 0; c < 3; ++c) {
        along_x(Ox, sm.g[c], sm.a[0], Kg, Kg, Kg, ox);
        __syncthreads();
        along_y(Oy, sm.a[0], sm.b[0], Kg, Kg, oy, ox);
        __syncthreads();
        along_z(Oz, sm.b[0], sm.fc[(op - 1) * 3 + c], Kg, oz, oy * ox);
        __syncthreads();
      }
    }
    const int L = a.nbr[(size_t) e * 6 + 2 * dir], R = a.nbr[(size_t) e * 6 + 2 * dir + 1];
    for (int l = tid; l < N2; l += kGeoThreads) {
      const double dr[3] = {sm.fc[0][l], sm.fc[1][l], sm.fc[2][l]};
      const double ds[3] = {sm.fc[3][l], sm.fc[4][l], sm.fc[5][l]};
      const double dt[3] = {sm.fc[6][l], sm.fc[7][l], sm.fc[8][l]};
      // normal along +r (I), +s (J), +t (K): ds x dt, dt x dr, dr x ds
      const double *u = dir == 0 ? ds : dir == 1 ? dt : dr;
      const double *v = dir == 0 ? dt : dir == 1 ? dr : ds;
      const double nx = u[1] * v[2] - u[2] * v[1];
      const double ny = u[2] * v[0] - u[0] * v[2];
      const double nz = u[0] * v[1] - u[1] * v[0];
      const double len = sqrt(nx * nx + ny * ny + nz * nz);
      double J, mi[9];
      inverse_metric(dr, ds, dt, J, mi);
      const double nrm[4] = {nx / len, ny / len, nz / len, len};
      for (int q = 0; q < 4; ++q) sm.fn[f][q][l] = nrm[q];
      // per-face slots of the face kernel: the min face of e is the right
      // side of slot (e, dir); its max face the left side of (R, dir)
      if (side == 0) {
        double *dst = a.fgeo + ((size_t) e * 3 + dir) * kFaceGeo * N2;
        for (int q = 0; q < 9; ++q) dst[(13 + q) * N2 + l] = mi[q];
        if (L < 0) {
          for (int q = 0; q < 4; ++q) dst[q * N2 + l] = nrm[q];
          for (int q = 0; q < 9; ++q) dst[(4 + q) * N2 + l] = mi[q];
        }
      } else if (R >= 0) {
        double *dst = a.fgeo + ((size_t) R * 3 + dir) * kFaceGeo * N2;
        for (int q = 0; q < 4; ++q) dst[q * N2 + l] = nrm[q];
        for (int q = 0; q < 9; ++q) dst[(4 + q) * N2 + l] = mi[q];
      }
    }
    __syncthreads();
  }
  // ---- element constants, sequential sums in the host set-up's order
  if (tid == 0) {
    double *ec = a.econst + (size_t) e * kElemConst;
    double vol = 0.0;
    for (int l = 0; l < NI; ++l) vol += w3[l] * sm.J[l];
    ec[3] = vol;
    for (int dir = 0; dir < 3; ++dir) {
      double an[3] = {0.0, 0.0, 0.0}, area = 0.0;
      for (int l = 0; l < N2; ++l) {
        const double w = w2[l];
        for (int c = 0; c < 3; ++c)
          an[c] += w * (sm.fn[2 * dir][c][l] * sm.fn[2 * dir][3][l] + sm.fn[2 * dir + 1][c][l] * sm.fn[2 * dir + 1][3][l]);
        area += w * (sm.fn[2 * dir][3][l] + sm.fn[2 * dir + 1][3][l]);
      }
      ec[dir] = 2.0 * vol / area;
      for (int c = 0; c < 3; ++c) ec[4 + 3 * dir + c] = an[c] / area;
    }
    for (int q = 13; q < kElemConst; ++q) ec[q] = 0.0;
  }
}

}  // namespace

hipError_t launch_geometry(const GeoArgs &a, int M, hipStream_t s)
{
  if (a.Kg < 2 || a.Kg > kMaxKg) return hipErrorInvalidValue;
  switch (M) {
    case 4: hipLaunchKernelGGL(geometry_kernel<4>, dim3(a.nElem), dim3(kGeoThreads), 0, s, a); break;
    case 5: hipLaunchKernelGGL(geometry_kernel<5>, dim3(a.nElem), dim3(kGeoThreads), 0, s, a); break;
    case 6: hipLaunchKernelGGL(geometry_kernel<6>, dim3(a.nElem), dim3(kGeoThreads), 0, s, a); break;
    case 7: hipLaunchKernelGGL(geometry_kernel<7>, dim3(a.nElem), dim3(kGeoThreads), 0, s, a); break;
    case 8: hipLaunchKernelGGL(geometry_kernel<8>, dim3(a.nElem), dim3(kGeoThreads), 0, s, a); break;
    case 9: hipLaunchKernelGGL(geometry_kernel<9>, dim3(a.nElem), dim3(kGeoThreads), 0, s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace vcp
