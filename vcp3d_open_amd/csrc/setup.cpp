// Host-side setup for the GPU path. Follows the reference setup code:
//   StandardElementClass::DataStandardElement   StandardElementClass.cpp:105-175
//   SolverClass::ComputeMetricTerms             SolverClass.cpp:829-1016
//   ElementClass::ComputeMetricTerms            ElementClass.cpp:1027-1407
//   ElementClass::FinalMetric{I,J,K}Face        ElementClass.cpp:1907-2088
//   ElementClass::ComputeLengthScales           ElementClass.cpp:927-1022
//   ElementClass::LGLDistribution               ElementClass.cpp:3406-3623
//   InputParamClass::ComputeReferenceConditions InputParamClass.cpp:605-758
//   StandardElementClass::AssemblePolynomialCorrectionOperators
//                                               StandardElementClass.cpp:176-322
//   SolverClass::InitNSCBC, ConfigureParamNSCBC SolverClass.cpp:2447-2497,
//                                               SubfaceClasses.cpp:1325-1384,
//                                               BoundaryState.cpp:829-925
//   SolverClass::InitSpongeLayer and friends    SpongeLayer.cpp:13-997
// (paths relative to VCP3D_CurviLinear/src/).
#include "setup.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <thread>
#include <stdexcept>

namespace vcp {

namespace {

// Orthonormal Legendre polynomial of degree n and its derivative at x
// (sqrt((2n+1)/2) P_n); the reference evaluates the same basis through
// normalised Jacobi recursions (StandAloneFunctions.cpp:96-138, :329-383).
void ortho_legendre(int n, double x, double &v, double &dv)
{
  double pm = 1.0, p = x, dpm = 0.0, dp = 1.0;
  if (n == 0) { p = 1.0; dp = 0.0; }
  for (int m = 1; m < n; ++m) {
    const double pn = ((2 * m + 1) * x * p - m * pm) / (m + 1);
    const double dpn = dpm + (2 * m + 1) * p;
    pm = p; p = pn; dpm = dp; dp = dpn;
  }
  const double s = std::sqrt(0.5 * (2 * n + 1));
  v = s * p;
  dv = s * dp;
}

// Plain Legendre P_n and P'_n (used for root finding).
void legendre_pn(int n, double x, double &p, double &dp)
{
  double pm = 1.0, pc = x;
  if (n == 0) { p = 1.0; dp = 0.0; return; }
  for (int m = 1; m < n; ++m) {
    const double pn = ((2 * m + 1) * x * pc - m * pm) / (m + 1);
    pm = pc; pc = pn;
  }
  p = pc;
  dp = n * (x * pc - pm) / (x * x - 1.0);
}

void gauss_points(int n, std::vector<double> &x, std::vector<double> &w)
{
  x.assign(n, 0.0); w.assign(n, 0.0);
  for (int i = 0; i < (n + 1) / 2; ++i) {
    double z = -std::cos(M_PI * (i + 0.75) / (n + 0.5));
    double p = 0, dp = 1;
    for (int it = 0; it < 100; ++it) {
      legendre_pn(n, z, p, dp);
      const double dz = p / dp;
      z -= dz;
      if (std::fabs(dz) < 1e-16) break;
    }
    legendre_pn(n, z, p, dp);
    const double wt = 2.0 / ((1.0 - z * z) * dp * dp);
    x[i] = z; x[n - 1 - i] = -z;
    w[i] = wt; w[n - 1 - i] = wt;
  }
  if (n % 2) x[n / 2] = 0.0;
}

// StandAloneFunctions.cpp:182-323 (LocationDOFs1D)
std::vector<double> edge_points(int n, int location)
{
  std::vector<double> r(n);
  if (location == 1) {
    for (int i = 0; i < n; ++i) r[i] = -1.0 + i * (2.0 / (n - 1));
    return r;
  }
  if (n == 1) { r[0] = 0.0; return r; }
  r[0] = -1.0; r[n - 1] = 1.0;
  if (n == 4) { r[1] = -std::sqrt(1.0 / 5.0); r[2] = -r[1]; return r; }
  if (n == 5) { r[1] = -std::sqrt(3.0 / 7.0); r[2] = 0.0; r[3] = -r[1]; return r; }
  if (n == 6) {
    const double t1 = 2.0 * std::sqrt(7.0) / 21.0;
    r[1] = -std::sqrt(1.0 / 3.0 + t1); r[2] = -std::sqrt(1.0 / 3.0 - t1); r[3] = -r[2]; r[4] = -r[1];
    return r;
  }
  if (n == 7) {
    const double t1 = 2.0 * std::sqrt(5.0 / 3.0) / 11.0;
    r[1] = -std::sqrt(5.0 / 11.0 + t1); r[2] = -std::sqrt(5.0 / 11.0 - t1); r[3] = 0.0; r[4] = -r[2]; r[5] = -r[1];
    return r;
  }
  // interior LGL points: roots of P'_{n-1}, Newton with (1-z^2)P'' = 2zP' - N(N+1)P
  const int N = n - 1;
  for (int i = 1; i <= (n - 1) / 2; ++i) {
    double z = -std::cos(M_PI * i / N);
    for (int it = 0; it < 100; ++it) {
      double p, dp;
      legendre_pn(N, z, p, dp);
      const double d2p = (2.0 * z * dp - N * (N + 1.0) * p) / (1.0 - z * z);
      const double dz = dp / d2p;
      z -= dz;
      if (std::fabs(dz) < 1e-16) break;
    }
    r[i] = z; r[n - 1 - i] = -z;
  }
  if (n % 2) r[n / 2] = 0.0;
  return r;
}

void lagrange_basis(const std::vector<double> &nodes, double x, double *l, double *dl)
{
  const int n = (int) nodes.size();
  for (int j = 0; j < n; ++j) {
    double v = 1.0, d = 0.0;
    for (int m = 0; m < n; ++m) {
      if (m == j) continue;
      const double f = 1.0 / (nodes[j] - nodes[m]);
      d = (d * (x - nodes[m]) + v) * f;
      v *= (x - nodes[m]) * f;
    }
    if (l) l[j] = v;
    if (dl) dl[j] = d;
  }
}

void invert_dense(int n, std::vector<double> &a)
{
  std::vector<double> b(n * n, 0.0);
  for (int i = 0; i < n; ++i) b[i * n + i] = 1.0;
  for (int c = 0; c < n; ++c) {
    int piv = c;
    for (int r = c + 1; r < n; ++r)
      if (std::fabs(a[r * n + c]) > std::fabs(a[piv * n + c])) piv = r;
    if (a[piv * n + c] == 0.0) throw std::runtime_error("singular Vandermonde matrix");
    for (int k = 0; k < n; ++k) { std::swap(a[c * n + k], a[piv * n + k]); std::swap(b[c * n + k], b[piv * n + k]); }
    const double inv = 1.0 / a[c * n + c];
    for (int k = 0; k < n; ++k) { a[c * n + k] *= inv; b[c * n + k] *= inv; }
    for (int r = 0; r < n; ++r) {
      if (r == c) continue;
      const double f = a[r * n + c];
      if (f == 0.0) continue;
      for (int k = 0; k < n; ++k) { a[r * n + k] -= f * a[c * n + k]; b[r * n + k] -= f * b[c * n + k]; }
    }
  }
  a.swap(b);
}

// Applies 1D operators (row-major out x in) along x, y and z of a tensor
// in[niz][niy][nix] -> out[noz][noy][nox].
void apply3(int nix, int niy, int niz, int nox, int noy, int noz,
            const double *ox, const double *oy, const double *oz, const double *in, double *out)
{
  std::vector<double> a((size_t) niz * niy * nox), b((size_t) niz * noy * nox);
  for (int k = 0; k < niz; ++k)
    for (int j = 0; j < niy; ++j)
      for (int i = 0; i < nox; ++i) {
        double s = 0.0;
        for (int q = 0; q < nix; ++q) s += ox[i * nix + q] * in[(k * niy + j) * nix + q];
        a[(k * niy + j) * nox + i] = s;
      }
  for (int k = 0; k < niz; ++k)
    for (int j = 0; j < noy; ++j)
      for (int i = 0; i < nox; ++i) {
        double s = 0.0;
        for (int q = 0; q < niy; ++q) s += oy[j * niy + q] * a[(k * niy + q) * nox + i];
        b[(k * noy + j) * nox + i] = s;
      }
  for (int k = 0; k < noz; ++k)
    for (int j = 0; j < noy; ++j)
      for (int i = 0; i < nox; ++i) {
        double s = 0.0;
        for (int q = 0; q < niz; ++q) s += oz[k * niz + q] * b[(q * noy + j) * nox + i];
        out[(k * noy + j) * nox + i] = s;
      }
}

// Classification of the grid-DOF spacing of one element; 1 LGL, 0 equidistant,
// -1 undecidable (the reference terminates in that case).
int classify_spacing(int Pg, const double *g, int nG)
{
  const int n = Pg + 1;
  const std::vector<double> req = edge_points(n, 1), rlgl = edge_points(n, 2);
  int nEq = 0, nLgl = 0, nBad = 0;
  std::vector<double> s(n);
  for (int dir = 0; dir < 3; ++dir)
    for (int a = 0; a < n; ++a)
      for (int b = 0; b < n; ++b) {
        s[0] = 0.0;
        for (int t = 1; t < n; ++t) {
          int i1 = 0, stride = 1;
          if (dir == 0) { i1 = (a * n + b) * n + t; stride = 1; }
          else if (dir == 1) { i1 = (a * n + t) * n + b; stride = n; }
          else { i1 = (t * n + a) * n + b; stride = n * n; }
          const int i0 = i1 - stride;
          double d2 = 0.0;
          for (int c = 0; c < 3; ++c) { const double d = g[c * nG + i1] - g[c * nG + i0]; d2 += d * d; }
          s[t] = s[t - 1] + std::sqrt(d2);
        }
        double devEq = 0.0, devL = 0.0;
        for (int t = 1; t < Pg; ++t) {
          const double v = -1.0 + 2.0 * s[t] / s[Pg];
          devEq = std::max(devEq, std::fabs(v - req[t]));
          devL = std::max(devL, std::fabs(v - rlgl[t]));
        }
        if (devEq <= devL) { ++nEq; nBad += devEq > 0.05; }
        else { ++nLgl; nBad += devL > 0.05; }
      }
  const int major = std::max(nEq, nLgl), minor = std::min(nEq, nLgl);
  if (major < 10 * minor || major < 10 * nBad) return -1;
  return nLgl > nEq ? 1 : 0;
}

void inverse_metric(const double dr[3], const double ds[3], const double dt[3], double &J, double m[9])
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

}  // namespace

int n_int1d(int rule) { return rule / 2 + 1; }

void build_tables(const CaseParams &p, Tables &t)
{
  t.P = p.n_poly_sol; t.K = t.P + 1; t.M = n_int1d(p.n_poly_int_rule); t.Kg = p.n_poly_grid + 1;
  const int K = t.K, M = t.M;
  gauss_points(M, t.r_int, t.w_int);
  t.r_dof = edge_points(K, p.dof_location);
  t.A.assign(M * K, 0); t.D.assign(M * K, 0); t.AT.assign(M * K, 0); t.DT.assign(M * K, 0);
  for (int pt = 0; pt < M; ++pt)
    for (int q = 0; q < K; ++q) {
      double v, d;
      ortho_legendre(q, t.r_int[pt], v, d);
      t.A[pt * K + q] = v; t.D[pt * K + q] = d;
      t.AT[q * M + pt] = v; t.DT[q * M + pt] = d;
    }
  t.LD.assign(K * K, 0); t.DLD.assign(K * K, 0);
  for (int pt = 0; pt < K; ++pt)
    for (int q = 0; q < K; ++q) {
      double v, d;
      ortho_legendre(q, t.r_dof[pt], v, d);
      t.LD[pt * K + q] = v; t.DLD[pt * K + q] = d;
    }
  t.VI = t.LD;
  invert_dense(K, t.VI);
  t.f_min.resize(K); t.f_max.resize(K); t.df_min.resize(K); t.df_max.resize(K);
  for (int q = 0; q < K; ++q) {
    ortho_legendre(q, -1.0, t.f_min[q], t.df_min[q]);
    ortho_legendre(q, 1.0, t.f_max[q], t.df_max[q]);
  }
  t.w3.resize(M * M * M); t.w2.resize(M * M);
  for (int k = 0; k < M; ++k)
    for (int j = 0; j < M; ++j) {
      t.w2[k * M + j] = t.w_int[j] * t.w_int[k];
      for (int i = 0; i < M; ++i) t.w3[(k * M + j) * M + i] = t.w_int[i] * t.w_int[j] * t.w_int[k];
    }
  std::vector<double> c0(K);
  for (int q = 0; q < K; ++q) { double d; ortho_legendre(q, 0.0, c0[q], d); }
  t.c0 = c0;
  t.basis_center.resize(K * K * K);
  for (int k = 0; k < K; ++k)
    for (int j = 0; j < K; ++j)
      for (int i = 0; i < K; ++i) t.basis_center[(k * K + j) * K + i] = c0[i] * c0[j] * c0[k];
  // Lagrange tables for the open-boundary and sponge terms
  t.Lq.assign(M * K, 0.0); t.dLq.assign(M * M, 0.0);
  for (int pt = 0; pt < M; ++pt) {
    lagrange_basis(t.r_dof, t.r_int[pt], &t.Lq[pt * K], nullptr);
    lagrange_basis(t.r_int, t.r_int[pt], nullptr, &t.dLq[pt * M]);
  }
  std::vector<double> dl(K);
  lagrange_basis(t.r_dof, t.r_dof[0], nullptr, dl.data());
  t.dell0 = -dl[0];
  // projector onto the polynomials of degree <= P sampled at the Gauss
  // points: Q Q^T with Q the (twice) Gram-Schmidt orthonormalised columns of
  // the Legendre table A. The 2D polynomial-correction matrix of the
  // reference is (P1 x P1) / dell0 (a Kronecker product of this projector).
  std::vector<double> Q(M * K);
  for (int q = 0; q < K; ++q) {
    for (int pt = 0; pt < M; ++pt) Q[pt * K + q] = t.A[pt * K + q];
    for (int pass = 0; pass < 2; ++pass)
      for (int r = 0; r < q; ++r) {
        double d = 0.0;
        for (int pt = 0; pt < M; ++pt) d += Q[pt * K + r] * Q[pt * K + q];
        for (int pt = 0; pt < M; ++pt) Q[pt * K + q] -= d * Q[pt * K + r];
      }
    double nrm = 0.0;
    for (int pt = 0; pt < M; ++pt) nrm += Q[pt * K + q] * Q[pt * K + q];
    nrm = 1.0 / std::sqrt(nrm);
    for (int pt = 0; pt < M; ++pt) Q[pt * K + q] *= nrm;
  }
  t.P1.assign(M * M, 0.0);
  for (int a = 0; a < M; ++a)
    for (int b = 0; b < M; ++b) {
      double v = 0.0;
      for (int q = 0; q < K; ++q) v += Q[a * K + q] * Q[b * K + q];
      t.P1[a * M + b] = v;
    }
}

std::string check_case(const CaseParams &p)
{
  for (int f = 0; f < 6; ++f) {
    const bool per = p.periodic[f / 2] != 0;
    if (per && p.bc[f] != kBCPeriodic) return "a periodic direction cannot carry a physical boundary";
    if (!per && (p.bc[f] < kBCOutflowNSCBC || p.bc[f] > kBCFarfield))
      return "unsupported boundary condition on a non-periodic block face";
  }
  if (p.outlet_average_local)
    return "\"Outlet use local (element) averaged Mach number\" is not supported";
  if (p.sgs_model < kSGSNone || p.sgs_model > kSGSVreman) return "Unknown \"Subgrid scale model\" specified";
  if (p.cml && !p.sponge_layer) return "a characteristic matching layer needs the sponge layer";
  if (p.sponge_layer) {
    if (p.n_poly_grid != p.n_poly_sol) return "sponge layer needs equal grid and solution degree";
    bool any = false;
    for (int f = 0; f < 6; ++f) {
      if (p.sponge_elems[f] < 0) return "Sponge layer cannot have negative number of elements";
      any |= p.sponge_elems[f] > 0;
    }
    if (!any) return "No sponge layer elements have been detected in a sponge layer simulation.";
  }
  return "";
}

// Lagrange tables of the grid DOFs: values and derivatives at the Gauss
// points [M][Kg], the face vectors (0/1) and derivatives at r = -1, +1 [Kg].
static void grid_lagrange(const Tables &t, std::vector<double> &Lg, std::vector<double> &dLg,
                          std::vector<double> lf[2], std::vector<double> dlf[2])
{
  const int M = t.M, Kg = t.Kg;
  Lg.assign(M * Kg, 0.0); dLg.assign(M * Kg, 0.0);
  for (int pt = 0; pt < M; ++pt) lagrange_basis(t.r_grid, t.r_int[pt], &Lg[pt * Kg], &dLg[pt * Kg]);
  for (int s = 0; s < 2; ++s) {
    lf[s].assign(Kg, 0.0); dlf[s].assign(Kg, 0.0);
    lf[s][s ? Kg - 1 : 0] = 1.0;
    lagrange_basis(t.r_grid, s ? 1.0 : -1.0, nullptr, dlf[s].data());
  }
}

std::vector<double> geometry_tables(const Tables &t)
{
  std::vector<double> Lg, dLg, lf[2], dlf[2];
  grid_lagrange(t, Lg, dLg, lf, dlf);
  std::vector<double> out(Lg);
  out.insert(out.end(), dLg.begin(), dLg.end());
  for (int s = 0; s < 2; ++s) out.insert(out.end(), lf[s].begin(), lf[s].end());
  for (int s = 0; s < 2; ++s) out.insert(out.end(), dlf[s].begin(), dlf[s].end());
  out.insert(out.end(), t.w2.begin(), t.w2.end());
  out.insert(out.end(), t.w3.begin(), t.w3.end());
  return out;
}

std::string build_geometry(const CaseParams &p, Tables &t, const double *grid, Geometry &g, bool all)
{
  const int ni = p.n_elem[0], nj = p.n_elem[1], nk = p.n_elem[2];
  const int nElem = ni * nj * nk, M = t.M, Kg = t.Kg, nG = Kg * Kg * Kg;
  const int nInt = M * M * M, n2 = M * M;
  const int nThreads = std::max(1, std::min(16, std::min((int) std::thread::hardware_concurrency(), nElem / 64)));
  // f(e0, e1) on contiguous element ranges, one per host thread
  auto in_parallel = [nThreads, nElem](auto &&f) {
    if (nThreads == 1) { f(0, nElem); return; }
    std::vector<std::thread> pool;
    for (int t = 0; t < nThreads; ++t)
      pool.emplace_back([&f, t, nThreads, nElem] { f((int) ((long) nElem * t / nThreads), (int) ((long) nElem * (t + 1) / nThreads)); });
    for (auto &th : pool) th.join();
  };
  // grid DOF spacing: all LGL or all equidistant
  std::vector<int> spacing(nElem);
  in_parallel([&](int e0, int e1) {
    for (int e = e0; e < e1; ++e) spacing[e] = classify_spacing(p.n_poly_grid, grid + (size_t) e * 3 * nG, nG);
  });
  int nL = 0, nE = 0;
  for (int e = 0; e < nElem; ++e) {
    if (spacing[e] < 0) return "Not able to determine the point distribution";
    (spacing[e] ? nL : nE)++;
  }
  if (nL && nE) return "Elements with both equidistant and LGL spacing encountered";
  t.r_grid = edge_points(Kg, nL ? 2 : 1);
  std::vector<double> Lg, dLg, lf[2], dlf[2];
  grid_lagrange(t, Lg, dLg, lf, dlf);
  // neighbours; a periodic block face aliases the opposite owned element
  g.nbr.assign((size_t) nElem * 6, -1);
  const int n[3] = {ni, nj, nk};
  for (int k = 0; k < nk; ++k)
    for (int j = 0; j < nj; ++j)
      for (int i = 0; i < ni; ++i) {
        const int e = i + ni * (j + nj * k);
        for (int d = 0; d < 3; ++d)
          for (int s = 0; s < 2; ++s) {
            int idx[3] = {i, j, k};
            idx[d] += s ? 1 : -1;
            if (idx[d] < 0 || idx[d] >= n[d]) {
              if (!p.periodic[d]) continue;
              idx[d] = (idx[d] + n[d]) % n[d];
            }
            g.nbr[(size_t) e * 6 + 2 * d + s] = idx[0] + ni * (idx[1] + nj * idx[2]);
          }
      }
  // elements whose host metric terms are needed: all, or (the device
  // computes the rest) those with a physical face (open-boundary set-up) or
  // in a sponge layer (damping functions at the point coordinates)
  std::vector<char> need(nElem, all ? 1 : 0);
  int nNeed = all ? nElem : 0;
  if (!all)
    for (int e = 0; e < nElem; ++e) {
      bool b = false;
      for (int f = 0; f < 6; ++f) b = b || g.nbr[(size_t) e * 6 + f] < 0;
      if (p.sponge_layer) {
        const int idx[3] = {e % ni, (e / ni) % nj, e / (ni * nj)};
        for (int d = 0; d < 3; ++d)
          b = b || idx[d] < p.sponge_elems[2 * d] || idx[d] >= n[d] - p.sponge_elems[2 * d + 1];
      }
      need[e] = b;
      nNeed += b;
    }
  g.vol.assign(nNeed ? (size_t) nElem * 13 * nInt : 0, 0.0);
  for (int f = 0; f < 6; ++f) g.face[f].assign(nNeed ? (size_t) nElem * 16 * n2 : 0, 0.0);
  g.lens.assign((size_t) nElem * 4, 0.0);
  g.len_les.assign((size_t) nElem, 0.0);
  g.ave_norm.assign((size_t) nElem * 9, 0.0);
  // metric terms at the nodal solution DOFs, for the velocity gradients of
  // the SGS model there (ElementClass::ComputeCoorNodalSolDOFs,
  // ElementClass.cpp:820-918)
  const int K = t.K, nD = K * K * K;
  const bool sgs = p.sgs_model != kSGSNone;
  std::vector<double> LgN(K * Kg), dLgN(K * Kg);
  for (int pt = 0; pt < K; ++pt) lagrange_basis(t.r_grid, t.r_dof[pt], &LgN[pt * Kg], &dLgN[pt * Kg]);
  g.nodal.assign(sgs ? (size_t) nElem * 10 * nD : 0, 0.0);
  // elements are independent: contiguous ranges on host threads, each with
  // its own scratch (the arithmetic per element is unchanged)
  std::atomic<int> badJacobian{0};
  auto elements = [&](int e0, int e1) {
  std::vector<double> X(3 * (size_t) nInt), DR(3 * (size_t) nInt), DS(3 * (size_t) nInt), DTt(3 * (size_t) nInt);
  std::vector<double> C(12 * (size_t) n2);
  std::vector<double> NR(sgs ? 3 * (size_t) nD : 0), NS(NR.size()), NT(NR.size());
  for (int e = e0; e < e1; ++e) {
    if (badJacobian.load(std::memory_order_relaxed)) return;
    const double *ge = grid + (size_t) e * 3 * nG;
    if (sgs) {
      for (int c = 0; c < 3; ++c) {
        apply3(Kg, Kg, Kg, K, K, K, dLgN.data(), LgN.data(), LgN.data(), ge + c * nG, &NR[c * nD]);
        apply3(Kg, Kg, Kg, K, K, K, LgN.data(), dLgN.data(), LgN.data(), ge + c * nG, &NS[c * nD]);
        apply3(Kg, Kg, Kg, K, K, K, LgN.data(), LgN.data(), dLgN.data(), ge + c * nG, &NT[c * nD]);
      }
      double *vn = &g.nodal[(size_t) e * 10 * nD];
      for (int l = 0; l < nD; ++l) {
        const double dr[3] = {NR[l], NR[nD + l], NR[2 * nD + l]};
        const double ds[3] = {NS[l], NS[nD + l], NS[2 * nD + l]};
        const double dt[3] = {NT[l], NT[nD + l], NT[2 * nD + l]};
        double J, m[9];
        inverse_metric(dr, ds, dt, J, m);
        if (!(J > 0.0)) { badJacobian.store(1); return; }
        vn[l] = J;
        for (int q = 0; q < 9; ++q) vn[(1 + q) * nD + l] = m[q];
      }
    }
    if (!need[e]) continue;
    double *vm = &g.vol[(size_t) e * 13 * nInt];
    for (int c = 0; c < 3; ++c) {
      apply3(Kg, Kg, Kg, M, M, M, Lg.data(), Lg.data(), Lg.data(), ge + c * nG, &X[c * nInt]);
      apply3(Kg, Kg, Kg, M, M, M, dLg.data(), Lg.data(), Lg.data(), ge + c * nG, &DR[c * nInt]);
      apply3(Kg, Kg, Kg, M, M, M, Lg.data(), dLg.data(), Lg.data(), ge + c * nG, &DS[c * nInt]);
      apply3(Kg, Kg, Kg, M, M, M, Lg.data(), Lg.data(), dLg.data(), ge + c * nG, &DTt[c * nInt]);
    }
    for (int l = 0; l < nInt; ++l) {
      const double dr[3] = {DR[l], DR[nInt + l], DR[2 * nInt + l]};
      const double ds[3] = {DS[l], DS[nInt + l], DS[2 * nInt + l]};
      const double dt[3] = {DTt[l], DTt[nInt + l], DTt[2 * nInt + l]};
      double J, m[9];
      inverse_metric(dr, ds, dt, J, m);
      if (!(J > 0.0)) { badJacobian.store(1); return; }
      vm[l] = J;
      for (int q = 0; q < 9; ++q) vm[(1 + q) * nInt + l] = m[q];
      for (int c = 0; c < 3; ++c) vm[(10 + c) * nInt + l] = X[c * nInt + l];
    }
    for (int f = 0; f < 6; ++f) {
      const int dir = f / 2, side = f % 2;
      const double *fv = lf[side].data(), *dfv = dlf[side].data();
      const double *Ox[4], *Oy[4], *Oz[4];
      int ox = M, oy = M, oz = M;
      // value, d/dr, d/ds, d/dt with the face vector in the normal slot
      for (int a = 0; a < 4; ++a) { Ox[a] = Lg.data(); Oy[a] = Lg.data(); Oz[a] = Lg.data(); }
      Ox[1] = dLg.data(); Oy[2] = dLg.data(); Oz[3] = dLg.data();
      if (dir == 0) { ox = 1; for (int a = 0; a < 4; ++a) Ox[a] = (a == 1) ? dfv : fv; }
      if (dir == 1) { oy = 1; for (int a = 0; a < 4; ++a) Oy[a] = (a == 2) ? dfv : fv; }
      if (dir == 2) { oz = 1; for (int a = 0; a < 4; ++a) Oz[a] = (a == 3) ? dfv : fv; }
      for (int a = 0; a < 4; ++a)
        for (int c = 0; c < 3; ++c)
          apply3(Kg, Kg, Kg, ox, oy, oz, Ox[a], Oy[a], Oz[a], ge + c * nG, &C[(a * 3 + c) * n2]);
      double *sm = &g.face[f][(size_t) e * 16 * n2];
      for (int l = 0; l < n2; ++l) {
        double dr[3], ds[3], dt[3];
        for (int c = 0; c < 3; ++c) { dr[c] = C[(3 + c) * n2 + l]; ds[c] = C[(6 + c) * n2 + l]; dt[c] = C[(9 + c) * n2 + l]; }
        // normal along +r (I), +s (J), +t (K): ds x dt, dt x dr, dr x ds
        const double *a = dir == 0 ? ds : dir == 1 ? dt : dr;
        const double *b = dir == 0 ? dt : dir == 1 ? dr : ds;
        const double nx = a[1] * b[2] - a[2] * b[1];
        const double ny = a[2] * b[0] - a[0] * b[2];
        const double nz = a[0] * b[1] - a[1] * b[0];
        const double len = std::sqrt(nx * nx + ny * ny + nz * nz);
        sm[0 * n2 + l] = nx / len; sm[1 * n2 + l] = ny / len; sm[2 * n2 + l] = nz / len; sm[3 * n2 + l] = len;
        double J, m[9];
        inverse_metric(dr, ds, dt, J, m);
        for (int q = 0; q < 9; ++q) sm[(4 + q) * n2 + l] = m[q];
        for (int c = 0; c < 3; ++c) sm[(13 + c) * n2 + l] = C[c * n2 + l];
      }
    }
    double vol = 0.0;
    for (int l = 0; l < nInt; ++l) vol += t.w3[l] * vm[l];
    g.lens[(size_t) e * 4 + 3] = vol;
    // ElementClass::ComputeLengthScales (ElementClass.cpp:937-940)
    g.len_les[e] = std::pow(vol, 1.0 / 3.0) / t.P;
    for (int dir = 0; dir < 3; ++dir) {
      const double *s0 = &g.face[2 * dir][(size_t) e * 16 * n2], *s1 = &g.face[2 * dir + 1][(size_t) e * 16 * n2];
      double an[3] = {0, 0, 0}, area = 0.0;
      for (int l = 0; l < n2; ++l) {
        const double w = t.w2[l];
        for (int c = 0; c < 3; ++c) an[c] += w * (s0[c * n2 + l] * s0[3 * n2 + l] + s1[c * n2 + l] * s1[3 * n2 + l]);
        area += w * (s0[3 * n2 + l] + s1[3 * n2 + l]);
      }
      g.lens[(size_t) e * 4 + dir] = 2.0 * vol / area;
      for (int c = 0; c < 3; ++c) g.ave_norm[(size_t) e * 9 + 3 * dir + c] = an[c] / area;
    }
  }
  };
  in_parallel(elements);
  if (badJacobian.load()) return "Negative or zero Jacobians encountered";
  return "";
}

// InputParamClass::ComputeReferenceConditions (InputParamClass.cpp:605-758):
// pressure, total pressure, density, temperature and total temperature from
// the prescribed data of the physical faces (a later face overriding an
// earlier one), else from the free-stream Mach number, Reynolds number and
// temperature.
RefConditions reference_conditions(const CaseParams &p)
{
  RefConditions r{};
  double pOut = -1.0, pTot = -1.0, rhoIn = -1.0, T = -1.0, TTot = -1.0;
  for (int f = 0; f < 6; ++f) {
    if (p.periodic[f / 2]) continue;
    const int bc = p.bc[f];
    if (bc == kBCOutflowNSCBC || bc == kBCOutflowSubsonic) pOut = p.bc_data[f][0];
    if (bc == kBCInflowStaticNSCBC) rhoIn = p.bc_data[f][0];
    if (bc == kBCInflowTotalNSCBC || bc == kBCInflowSubsonic) { pTot = p.bc_data[f][0]; TTot = p.bc_data[f][1]; }
    if (bc == kBCWallIsothermal) T = p.bc_data[f][0];
    if (bc == kBCInflowSupersonic) { rhoIn = p.bc_data[f][0]; pOut = p.bc_data[f][4]; }
  }
  r.p_ref = pOut > 0.0 ? pOut : pTot > 0.0 ? pTot : -1.0;
  r.rho_ref = rhoIn > 0.0 ? rhoIn : -1.0;
  if (r.rho_ref <= 0.0 && r.p_ref > 0.0) {
    const double TRef = T > 0.0 ? T : TTot;
    if (TRef > 0.0) r.rho_ref = r.p_ref / (kRGas * TRef);
  }
  r.fallback = 0;
  if (r.p_ref <= 0.0 || r.rho_ref <= 0.0) {
    r.fallback = 1;
    const double vInf = p.mach * std::sqrt(kGamma * kRGas * p.temp_free_stream);
    r.rho_ref = p.reynolds * kMuDim / (p.reynolds_length * vInf);
    r.p_ref = r.rho_ref * kRGas * p.temp_free_stream;
  }
  r.u_ref = std::sqrt(r.p_ref / r.rho_ref);
  r.mu = kMuDim * r.u_ref / r.p_ref;
  return r;
}

// InputParamClass::DetermineNonDimPrimVarFreeStream (InputParamClass.cpp:
// 991-1191): the free stream from the prescribed data when they fix it
// (velocity and two of p, rho, T; or a velocity direction and two total
// conditions with the Mach number from a static/total pair), else from the
// Mach number, temperature, direction (unit when the reference conditions
// came from the same data, InputParamClass.cpp:697-707) and Reynolds number;
// non-dimensional (rho, u, v, w, p).
static void free_stream_nondim(const CaseParams &p, const RefConditions &rc, double q[5])
{
  const double big = 1e300;
  double pr = -1.0, pTot = -1.0, rho = -1.0, T = -1.0, TTot = -1.0;
  double vel[3] = {big, big, big}, dir[3] = {big, big, big};
  for (int f = 0; f < 6; ++f) {
    if (p.periodic[f / 2]) continue;
    const int bc = p.bc[f];
    const double *d = p.bc_data[f];
    if (bc == kBCOutflowNSCBC || bc == kBCOutflowSubsonic) pr = d[0];
    if (bc == kBCInflowStaticNSCBC) { rho = d[0]; vel[0] = d[1]; vel[1] = d[2]; vel[2] = d[3]; }
    if (bc == kBCInflowTotalNSCBC || bc == kBCInflowSubsonic) {
      pTot = d[0]; TTot = d[1]; dir[0] = d[2]; dir[1] = d[3]; dir[2] = d[4];
    }
    if (bc == kBCInflowSupersonic) { rho = d[0]; vel[0] = d[1]; vel[1] = d[2]; vel[2] = d[3]; pr = d[4]; }
    if (bc == kBCWallIsothermal) T = d[0];
  }
  const double gm1 = kGamma - 1.0;
  bool found = false;
  if (vel[0] < big && vel[1] < big && vel[2] < big) {
    const int n = (pr > 0.0) + (rho > 0.0) + (T > 0.0);
    if (n >= 2) {
      if (pr <= 0.0) pr = rho * kRGas * T;
      if (rho <= 0.0) rho = pr / (kRGas * T);
      found = true;
    } else if (n == 1 && T <= 0.0) {
      const double vm2 = vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2];
      if (pr <= 0.0) pr = 0.95 * rho * vm2 / kGamma; else rho = kGamma * pr / (0.95 * vm2);
      found = true;
    }
    if (found) { q[0] = rho; q[1] = vel[0]; q[2] = vel[1]; q[3] = vel[2]; q[4] = pr; }
  } else if (dir[0] < big && dir[1] < big && dir[2] < big) {
    const double inv = 1.0 / std::sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
    double M2 = 0.95 * 0.95;
    if (pTot > 0.0 && pr > 0.0) M2 = (std::pow(pTot / pr, gm1 / kGamma) - 1.0) * 2.0 / gm1;
    else if (TTot > 0.0 && T > 0.0) M2 = (TTot / T - 1.0) * 2.0 / gm1;
    if ((pTot > 0.0) + (TTot > 0.0) >= 2) {
      const double rhoTot = pTot / (kRGas * TTot);
      const double tmp = 1.0 + 0.5 * gm1 * M2;
      pr = pTot / std::pow(tmp, kGamma / gm1);
      rho = rhoTot / std::pow(tmp, 1.0 / gm1);
      const double vm = std::sqrt(M2 * kGamma * pr / rho);
      q[0] = rho; q[1] = dir[0] * inv * vm; q[2] = dir[1] * inv * vm; q[3] = dir[2] * inv * vm; q[4] = pr;
      found = true;
    }
  }
  if (!found) {
    double dv[3] = {p.free_stream_dir[0], p.free_stream_dir[1], p.free_stream_dir[2]};
    if (rc.fallback) {
      const double inv = 1.0 / std::sqrt(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
      for (double &x : dv) x *= inv;
    }
    const double vInf = p.mach * std::sqrt(kGamma * kRGas * p.temp_free_stream);
    const double muDim = rc.mu * rc.p_ref / rc.u_ref;
    const double rhoInf = p.reynolds * muDim / (p.reynolds_length * vInf);
    q[0] = rhoInf; q[1] = vInf * dv[0]; q[2] = vInf * dv[1]; q[3] = vInf * dv[2];
    q[4] = rhoInf * kRGas * p.temp_free_stream;
  }
  q[0] /= rc.rho_ref; q[1] /= rc.u_ref; q[2] /= rc.u_ref; q[3] /= rc.u_ref; q[4] /= rc.p_ref;
}

std::string setup_open_boundary(const CaseParams &p, const Tables &t, const Geometry &g,
                                const RefConditions &rc, OpenBoundary &ob)
{
  const int nElem = p.n_elem[0] * p.n_elem[1] * p.n_elem[2], n2 = t.M * t.M;
  ob = OpenBoundary();
  const double TRef = rc.p_ref / rc.rho_ref;
  for (int f = 0; f < 6; ++f) {
    ob.begin[f] = (int) ob.faces.size() / 2;
    if (p.periodic[f / 2]) continue;
    ob.any = true;
    const double *d = p.bc_data[f];
    double *o = ob.data[f];
    const int bc = p.bc[f];
    // ConvertPrescribedData (SubfaceClasses.cpp:739-751, :863-915,
    // :1169-1180, :1304-1313, :1385-1400, :1453-1475); an isothermal wall
    // keeps its dimensional temperature
    if (bc == kBCOutflowNSCBC || bc == kBCOutflowSubsonic) {
      o[0] = d[0] / rc.p_ref;
    } else if (bc == kBCInflowStaticNSCBC) {
      o[0] = d[0] / rc.rho_ref;
      for (int c = 1; c < 4; ++c) o[c] = d[c] / rc.u_ref;
    } else if (bc == kBCInflowTotalNSCBC || bc == kBCInflowSubsonic) {
      const double len = std::sqrt(d[2] * d[2] + d[3] * d[3] + d[4] * d[4]);
      if (!(len > 0.0)) return "total-condition inflow needs a velocity direction";
      o[0] = d[0] / rc.p_ref;
      o[1] = bc == kBCInflowSubsonic ? d[1] * (kGamma * kRGas / (kGamma - 1.0) / (rc.u_ref * rc.u_ref)) : d[1] / TRef;
      for (int c = 2; c < 5; ++c) o[c] = d[c] / len;
    } else if (bc == kBCWallHeatFlux) {
      o[0] = d[0] / (rc.p_ref * rc.u_ref);
    } else if (bc == kBCWallIsothermal) {
      if (!(d[0] > 0.0)) return "isothermal wall needs a positive temperature";
      o[0] = d[0];
      o[1] = -rc.p_ref / (kRGas * rc.rho_ref);   // the wall state's V4 is o[1] / o[0]
    } else if (bc == kBCInflowSupersonic) {
      // the prescribed state as entropy variables (SubfaceClasses.cpp:1026-1106)
      const double rho = d[0] / rc.rho_ref, u = d[1] / rc.u_ref, v = d[2] / rc.u_ref, w = d[3] / rc.u_ref;
      const double pr = d[4] / rc.p_ref;
      const double sEnt = std::log(pr / std::pow(rho, kGamma)), pInv = 1.0 / pr;
      o[0] = (kGamma - sEnt) / (kGamma - 1.0) - 0.5 * pInv * rho * (u * u + v * v + w * w);
      o[1] = rho * u * pInv; o[2] = rho * v * pInv; o[3] = rho * w * pInv; o[4] = -rho * pInv;
    } else if (bc == kBCFarfield) {
      // the free stream as entropy variables (SubfaceClasses.cpp:542-622)
      double q[5];
      free_stream_nondim(p, rc, q);
      const double sEnt = std::log(q[4] / std::pow(q[0], kGamma)), pInv = 1.0 / q[4];
      o[0] = (kGamma - sEnt) / (kGamma - 1.0) - 0.5 * pInv * q[0] * (q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
      o[1] = q[0] * q[1] * pInv; o[2] = q[0] * q[2] * pInv; o[3] = q[0] * q[3] * pInv; o[4] = -q[0] * pInv;
    } else if (bc != kBCWallInviscid && bc != kBCOutflowSupersonic) {
      return "unsupported boundary condition";
    }
    int first = -1;
    for (int e = 0; e < nElem; ++e) {
      if (g.nbr[(size_t) e * 6 + f] >= 0) continue;
      if (first < 0) first = e;
      ob.faces.push_back(e);
      ob.faces.push_back(f);
      const double *sm = &g.face[f][(size_t) e * 16 * n2];
      for (int l = 0; l < n2; ++l) ob.area[f] += t.w2[l] * sm[3 * n2 + l];
    }
    // orientation: the parametric direction whose gradient is parallel to
    // the face normal (integration point 0 of the first boundary element)
    const double *sm = &g.face[f][(size_t) first * 16 * n2];
    int normal = -1;
    for (int i = 0; i < 3; ++i) {
      const double *gr = sm + (size_t) (4 + 3 * i) * n2;
      const double len = std::sqrt(gr[0] * gr[0] + gr[n2] * gr[n2] + gr[2 * n2] * gr[2 * n2]);
      const double ct = (sm[0] * gr[0] + sm[n2] * gr[n2] + sm[2 * n2] * gr[2 * n2]) / len;
      if (std::fabs(ct - 1.0) < 1e-10) normal = i;
    }
    if (!bc_is_nscbc(bc)) continue;   // the orientation serves the characteristic faces only
    if (normal < 0) return "Could not identify relevant boundary indices!";
    if (normal != f / 2) return "NSCBC face whose normal is not along its own parametric direction";
    ob.ig[f][0] = normal;
    ob.ig[f][1] = normal == 0 ? 1 : 0;
    ob.ig[f][2] = normal == 2 ? 1 : 2;
  }
  ob.begin[6] = (int) ob.faces.size() / 2;
  return "";
}

std::string setup_sponge(const CaseParams &p, const Tables &t, const Geometry &g, const double *grid,
                         Sponge &sp)
{
  sp = Sponge();
  if (!p.sponge_layer) return "";
  const int ni = p.n_elem[0], nj = p.n_elem[1], nk = p.n_elem[2], nElem = ni * nj * nk;
  const int n[3] = {ni, nj, nk};
  const int K = t.K, nInt = t.M * t.M * t.M, nG = t.Kg * t.Kg * t.Kg;
  std::vector<int> flag(nElem, 0);
  sp.idx.assign(nElem, -1);
  for (int e = 0; e < nElem; ++e) {
    const int idx[3] = {e % ni, (e / ni) % nj, e / (ni * nj)};
    for (int d = 0; d < 3; ++d) {
      if (idx[d] < p.sponge_elems[2 * d]) flag[e] |= 1 << (2 * d);
      if (idx[d] >= n[d] - p.sponge_elems[2 * d + 1]) flag[e] |= 1 << (2 * d + 1);
    }
    if (flag[e]) sp.idx[e] = sp.n++;
  }
  sp.fn.assign((size_t) sp.n * 3 * nInt, 0.0);
  sp.coef.assign((size_t) sp.n * 8, 0.0);
  for (int e = 0; e < nElem; ++e)
    if (sp.idx[e] >= 0)
      for (int d = 0; d < 3; ++d) sp.coef[(size_t) sp.idx[e] * 8 + 3 + d] = (flag[e] >> (2 * d)) & 1 ? -1.0 : 1.0;
  for (int f = 0; f < 6; ++f) {
    if (p.sponge_elems[f] <= 0) continue;
    const int d = f / 2, bit = 1 << f;
    const int last = p.n_poly_sol * (d == 0 ? 1 : d == 1 ? K : K * K);   // node index P along d
    double lo = 1e300, hi = -1e300;
    for (int e = 0; e < nElem; ++e) {
      if (!(flag[e] & bit)) continue;
      const double *ge = grid + ((size_t) e * 3 + d) * nG;
      lo = std::min(lo, ge[0]);
      hi = std::max(hi, ge[last]);
    }
    const double inv = 1.0 / (hi - lo), x0 = (f % 2) ? lo : hi;
    const double sig = p.damping_constant[f], alpha = p.damping_exponent[f];
    for (int e = 0; e < nElem; ++e) {
      if (!(flag[e] & bit)) continue;
      const double *x = &g.vol[((size_t) e * 13 + 10 + d) * nInt];
      double *fn = &sp.fn[((size_t) sp.idx[e] * 3 + d) * nInt];
      for (int l = 0; l < nInt; ++l) {
        if ((f % 2) ? x[l] < x0 : x[l] > x0) return "Inconsistent sponge layer location.";
        fn[l] = sig * std::pow(inv * std::fabs(x[l] - x0), alpha);
      }
      if (p.cml) sp.coef[(size_t) sp.idx[e] * 8 + d] = std::fabs(sig) < 1.0e-4 ? 1.0 : 1.0 / sig;
    }
  }
  return "";
}

void damping_state(const Tables &t, const RefConditions &rc, const double *nodal_prim, double *out)
{
  const int K = t.K, M = t.M, nD = K * K * K, nInt = M * M * M;
  std::vector<double> node((size_t) kNVar * nD), q((size_t) kNVar * nInt);
  for (int d = 0; d < nD; ++d)
    for (int l = 0; l < kNVar; ++l) node[(size_t) l * nD + d] = nodal_prim[(size_t) d * kNVar + l];
  for (int l = 0; l < kNVar; ++l)
    apply3(K, K, K, M, M, M, t.Lq.data(), t.Lq.data(), t.Lq.data(), &node[(size_t) l * nD], &q[(size_t) l * nInt]);
  for (int m = 0; m < nInt; ++m) {
    const double rho = q[m] / rc.rho_ref, u = q[nInt + m] / rc.u_ref, v = q[2 * nInt + m] / rc.u_ref,
                 w = q[3 * nInt + m] / rc.u_ref, pr = q[4 * nInt + m] / rc.p_ref;
    out[m] = rho;
    out[nInt + m] = rho * u;
    out[2 * nInt + m] = rho * v;
    out[3 * nInt + m] = rho * w;
    out[4 * nInt + m] = pr / (kGamma - 1.0) + 0.5 * rho * (u * u + v * v + w * w);
  }
}

void nodal_cons_to_modal_entropy(const Tables &t, const RefConditions &rc, const double *nodal, double *modal)
{
  const int K = t.K, nD = K * K * K;
  const double gm1 = kGamma - 1.0;
  std::vector<double> v((size_t) kNVar * nD);
  for (int d = 0; d < nD; ++d) {
    const double *s = nodal + (size_t) d * kNVar;
    const double ri = 1.0 / s[0];
    const double u = ri * s[1], vv = ri * s[2], w = ri * s[3];
    const double p = gm1 * (s[4] - 0.5 * (u * s[1] + vv * s[2] + w * s[3])) / rc.p_ref;
    const double rho = s[0] / rc.rho_ref, un = u / rc.u_ref, vn = vv / rc.u_ref, wn = w / rc.u_ref;
    const double ent = std::log(p / std::pow(rho, kGamma)), pi = 1.0 / p;
    v[0 * nD + d] = (kGamma - ent) / gm1 - 0.5 * rho * (un * un + vn * vn + wn * wn) * pi;
    v[1 * nD + d] = rho * un * pi;
    v[2 * nD + d] = rho * vn * pi;
    v[3 * nD + d] = rho * wn * pi;
    v[4 * nD + d] = -rho * pi;
  }
  for (int l = 0; l < kNVar; ++l)
    apply3(K, K, K, K, K, K, t.VI.data(), t.VI.data(), t.VI.data(), &v[(size_t) l * nD], modal + (size_t) l * nD);
}

void modal_entropy_to_nodal_cons(const Tables &t, const RefConditions &rc, const double *modal, double *nodal)
{
  const int K = t.K, nD = K * K * K;
  const double gm1 = kGamma - 1.0;
  std::vector<double> v((size_t) kNVar * nD);
  for (int l = 0; l < kNVar; ++l)
    apply3(K, K, K, K, K, K, t.LD.data(), t.LD.data(), t.LD.data(), modal + (size_t) l * nD, &v[(size_t) l * nD]);
  const double momRef = rc.rho_ref * rc.u_ref;
  for (int d = 0; d < nD; ++d) {
    const double V4i = 1.0 / v[4 * nD + d];
    const double u = -V4i * v[1 * nD + d], vv = -V4i * v[2 * nD + d], w = -V4i * v[3 * nD + d];
    const double ek = 0.5 * (u * u + vv * vv + w * w);
    const double s = kGamma - gm1 * (v[0 * nD + d] - v[4 * nD + d] * ek);
    const double rho = std::pow(-v[4 * nD + d] * std::exp(s), -1.0 / gm1);
    const double p = -rho * V4i;
    double *o = nodal + (size_t) d * kNVar;
    o[0] = rc.rho_ref * rho;
    o[1] = momRef * rho * u; o[2] = momRef * rho * vv; o[3] = momRef * rho * w;
    o[4] = rc.p_ref * (rho * ek + p / gm1);
  }
}

}  // namespace vcp
