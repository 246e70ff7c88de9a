/*
 * vcp_oracle.c -- TEST INFRASTRUCTURE ONLY (the checker, never the product).
 *
 * Plain-C restatement of the VCP3D explicit DG residual and SSP-RK3
 * synchronization step for entropy variables + Ismail-Roe, periodic blocks.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this library. Every routine cites the reference file:line it follows
 * (paths relative to VCP3D_CurviLinear/).
 *
 * Pinning (what fixes each function besides reading the reference):
 *   tables      -- Gauss-Legendre vs numpy.leggauss, orthonormality, LGL
 *                  closed forms, Lagrange partition of unity (tests).
 *   tensor prod -- dense Kronecker products (tests).
 *   metrics     -- affine grids: constant J = (h/2)^3, exact normals.
 *   residual    -- reference golden vectors (oracle/_ref build), free-stream
 *                  preservation, conservation, TGV symmetry (tests).
 *   CG          -- numpy.linalg.solve of the assembled mass matrix (tests).
 *   RK step     -- reference golden trajectories.
 *   SGS models  -- closed forms (solid-body rotation, axisymmetric strain),
 *                  zero in pure shear / isotropic expansion, frame invariance
 *                  and homogeneity (tests); reference golden vectors of WALE
 *                  and Vreman cases, including the monitored maximum.
 */
#define _USE_MATH_DEFINES
#include "vcp_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#define NVAR 5
static const double GAM = 1.4;
static const double PRANDTL_LAM = 0.72;              /* VCP3D_CurviLinear.hpp:175 */
static const double PRANDTL_TURB = 0.90;             /* VCP3D_CurviLinear.hpp:176 */
static const double EPS_SMALL = 1.e-30;              /* VCP3D_CurviLinear.hpp:140 */
static const double RGAS = 287.058;                  /* VCP3D_CurviLinear.hpp:177 */
static const double LAMBDA_OVER_MU = -2.0 / 3.0;     /* VCP3D_CurviLinear.hpp:186 */

static char g_err[512];
const char *orc_last_error(void) { return g_err; }
static int fail(const char *msg) { snprintf(g_err, sizeof g_err, "%s", msg); return -1; }

struct orc_case {
  orc_params prm;
  int ni, nj, nk, nElem;
  int P, K, M, Kg, nDOFs, nInt, nInt2D, nGrid;
  double pRef, uRef, rhoRef, mu;
  /* 1D tables, operators stored row-major as (n_out x n_in) */
  double *rInt, *wInt, *rDOF, *rGrid;
  double *A, *D;          /* M x K : phi_dof(r_int), phi'_dof(r_int)  */
  double *AT, *DT;        /* K x M                                  */
  double *LD;             /* K x K : phi_mode(r_nodal)  (modal->nodal) */
  double *DLD;            /* K x K : phi'_mode(r_nodal) (modal->nodal derivative) */
  double *VI;             /* K x K : inverse of LD       (nodal->modal) */
  double *FMin, *FMax, *dFMin, *dFMax; /* 1 x K face vectors (also K x 1) */
  double *w3, *w2, *basisCenter;
  double *Lq;             /* M x K : nodal DOF Lagrange basis at the Gauss points */
  double *dLq;            /* M x M : Gauss-point Lagrange derivative, [point][basis] */
  double *PC;             /* nInt2D x nInt2D polynomial-correction matrix */
  double dell0;           /* -l'_0(r_0) of the nodal DOF basis */
  /* physical boundaries (PC-NSCBC) */
  int nscbc;              /* any NSCBC face */
  int ig[6][3];           /* gradient slots (normal, first, second transverse) */
  double bcData[6][5];    /* non-dimensional prescribed data */
  double faceArea[6], machAvg[6];
  /* sponge layer / characteristic matching layer */
  int *spongeIdx;         /* [nElem] -> sponge slot or -1 */
  int nSponge;
  unsigned char *spongeFlag; /* bit 2*dir+side: element lies in that layer */
  double *dampFn;         /* [nSponge][3][nInt] */
  double *dampState;      /* [nSponge][5][nInt] conservative, non-dimensional */
  double *matchScale;     /* [nSponge][3] normalisation of the damping function */
  int dampStateSet;
  /* geometry */
  double *vm;             /* [nElem][13][nInt] */
  double *sm[6];          /* [nElem][16][nInt2D] iMin,iMax,jMin,jMax,kMin,kMax */
  double *len;            /* [nElem][4]: I, J, K, volume */
  double *lenLES;         /* [nElem]: volume^(1/3) / P, the LES length scale */
  double *vmN;            /* [nElem][10][nDOFs]: J, dr/dx .. dt/dz at the nodal
                             DOFs (SGS model only) */
  double *aveNorm;        /* [nElem][3][3] */
  int *nbr;               /* [nElem][6] */
  /* state */
  double *sol, *solOld, *res, *resMin[3], *dUdV;
  int *cgIt;
  int resValid, timeStep;
  int monNow;             /* the residual being evaluated computes monitoring data */
  double *forceElem;      /* [nElem][6] wall force sums of the element's faces */
  double *eddyElem;       /* [nElem] eddy-viscosity maximum over the element's faces */
  double *avg;            /* [nElem][nDOFs][12] running averages (compute_average) */
  int refFallback;        /* pRef, rhoRef from the free-stream data of the parameter file */
  int nAvg;
  /* per-thread scratch */
  double **works;
  int nthreads;
  size_t work_size;
};

static double *W(const orc_case *c) { return c->works[omp_get_thread_num()]; }
static int is_nscbc(int bc);
static void nscbc_average(orc_case *c);

/* ------------------------------------------------------------------ utils */
static double *dalloc(size_t n) { return (double *) calloc(n ? n : 1, sizeof(double)); }

/* out[k][j][i] (+)= sum Ox[i][ii] Oy[j][jj] Oz[k][kk] in[kk][jj][ii]; the
   sum-factorisation of TPSAGV/TPAVR/TPSAGI/TPRIF (include/TensorProduct*.hpp)
   written as three 1D operator applications. */
static void tp3(int nix, int niy, int niz, int nox, int noy, int noz,
                const double *ox, const double *oy, const double *oz,
                const double *in, double *out, int acc, double *w)
{
  double *t1 = w;                          /* [niz][niy][nox] */
  double *t2 = w + (size_t) niz * niy * nox; /* [niz][noy][nox] */
  for (int k = 0; k < niz; ++k)
    for (int j = 0; j < niy; ++j)
      for (int i = 0; i < nox; ++i) {
        double s = 0.0;
        for (int ii = 0; ii < nix; ++ii) s += ox[i * nix + ii] * in[(k * niy + j) * nix + ii];
        t1[(k * niy + j) * nox + i] = s;
      }
  for (int k = 0; k < niz; ++k)
    for (int j = 0; j < noy; ++j)
      for (int i = 0; i < nox; ++i) {
        double s = 0.0;
        for (int jj = 0; jj < niy; ++jj) s += oy[j * niy + jj] * t1[(k * niy + jj) * nox + i];
        t2[(k * noy + j) * nox + i] = s;
      }
  for (int k = 0; k < noz; ++k)
    for (int j = 0; j < noy; ++j)
      for (int i = 0; i < nox; ++i) {
        double s = 0.0;
        for (int kk = 0; kk < niz; ++kk) s += oz[k * niz + kk] * t2[(kk * noy + j) * nox + i];
        if (acc) out[(k * noy + j) * nox + i] += s; else out[(k * noy + j) * nox + i] = s;
      }
}

/* --------------------------------------------------------- 1D polynomials */
/* Orthonormal Legendre phi_n = sqrt((2n+1)/2) P_n and its derivative; same
   basis as NormJacobi(n,0,0,x) / GradNormJacobi (StandAloneFunctions.cpp:96-138,
   :329-383). */
static void legendre(int n, double x, double *p, double *dp)
{
  double p0 = 1.0, p1 = x, d0 = 0.0, d1 = 1.0;
  if (n == 0) { p1 = 1.0; d1 = 0.0; }
  for (int m = 1; m < n; ++m) {
    const double p2 = ((2.0 * m + 1.0) * x * p1 - m * p0) / (m + 1.0);
    const double d2 = d0 + (2.0 * m + 1.0) * p1;
    p0 = p1; p1 = p2; d0 = d1; d1 = d2;
  }
  const double s = sqrt((2.0 * n + 1.0) / 2.0);
  *p = s * p1; *dp = s * d1;
}

/* Gauss-Legendre nodes (ascending) and weights by Newton iteration; the
   reference obtains the same rule from Golub-Welsch with alpha=beta=0
   (StandardElementClass.cpp:662-689, GaussJacobiQuadrature.cpp). */
static void gauss_legendre(int n, double *x, double *w)
{
  for (int i = 0; i < n; ++i) {
    double z = cos(M_PI * (i + 0.75) / (n + 0.5));
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = z;
      for (int m = 1; m < n; ++m) { double p2 = ((2.0 * m + 1.0) * z * p1 - m * p0) / (m + 1.0); p0 = p1; p1 = p2; }
      const double pn = (n == 0) ? 1.0 : p1, pnm1 = p0;
      const double dp = n * (z * pn - pnm1) / (z * z - 1.0);
      const double dz = pn / dp;
      z -= dz;
      if (fabs(dz) < 1e-16) break;
    }
    double p0 = 1.0, p1 = z;
    for (int m = 1; m < n; ++m) { double p2 = ((2.0 * m + 1.0) * z * p1 - m * p0) / (m + 1.0); p0 = p1; p1 = p2; }
    const double dp = n * (z * p1 - p0) / (z * z - 1.0);
    x[n - 1 - i] = z;
    w[n - 1 - i] = 2.0 / ((1.0 - z * z) * dp * dp);
  }
  if (n % 2 == 1) x[n / 2] = 0.0;
}

/* DOF locations in the standard edge (StandAloneFunctions.cpp:182-323). */
static int dof_locations(int n, int loc, double *r)
{
  if (loc == 1) {
    const double dh = 2.0 / (n - 1);
    for (int i = 0; i < n; ++i) r[i] = -1.0 + i * dh;
    return 0;
  }
  switch (n) {
    case 1: r[0] = 0.0; break;
    case 2: r[0] = -1.0; r[1] = 1.0; break;
    case 3: r[0] = -1.0; r[1] = 0.0; r[2] = 1.0; break;
    case 4: r[0] = -1.0; r[1] = -sqrt(1.0 / 5.0); r[2] = sqrt(1.0 / 5.0); r[3] = 1.0; break;
    case 5: r[0] = -1.0; r[1] = -sqrt(3.0 / 7.0); r[2] = 0.0; r[3] = sqrt(3.0 / 7.0); r[4] = 1.0; break;
    case 6: { const double t1 = 2.0 * sqrt(7.0) / 21.0, t2 = sqrt(1.0 / 3.0 + t1), t3 = sqrt(1.0 / 3.0 - t1);
              r[0] = -1.0; r[1] = -t2; r[2] = -t3; r[3] = t3; r[4] = t2; r[5] = 1.0; break; }
    case 7: { const double t1 = 2.0 * sqrt(5.0 / 3.0) / 11.0, t2 = sqrt(5.0 / 11.0 + t1), t3 = sqrt(5.0 / 11.0 - t1);
              r[0] = -1.0; r[1] = -t2; r[2] = -t3; r[3] = 0.0; r[4] = t3; r[5] = t2; r[6] = 1.0; break; }
    default: {
      /* roots of P'_{n-1} by Newton (LGL), symmetric */
      r[0] = -1.0; r[n - 1] = 1.0;
      for (int i = 1; i < n - 1; ++i) {
        double z = -cos(M_PI * i / (n - 1));
        for (int it = 0; it < 100; ++it) {
          double p, dp, pm, dpm;
          legendre(n - 1, z, &p, &dp);
          legendre(n - 2, z, &pm, &dpm);
          /* (1-z^2) P'' = 2 z P' - N(N+1) P for P = P_{N}, N = n-1 */
          const double s = sqrt((2.0 * (n - 1) + 1.0) / 2.0);
          const double P1 = dp / s, P0 = p / s;
          const double P2 = (2.0 * z * P1 - (n - 1.0) * n * P0) / (1.0 - z * z);
          const double dz = P1 / P2;
          z -= dz;
          if (fabs(dz) < 1e-16) break;
        }
        r[i] = z;
      }
      for (int i = 1; i < n / 2; ++i) { const double a = 0.5 * (r[n - 1 - i] - r[i]); r[i] = -a; r[n - 1 - i] = a; }
      if (n % 2 == 1) r[n / 2] = 0.0;
    }
  }
  return 0;
}

/* Lagrange basis l_j of nodes xn at x (value and derivative). */
static void lagrange(int n, const double *xn, double x, double *l, double *dl)
{
  for (int j = 0; j < n; ++j) {
    double v = 1.0, d = 0.0;
    for (int m = 0; m < n; ++m) {
      if (m == j) continue;
      const double f = 1.0 / (xn[j] - xn[m]);
      d = d * (x - xn[m]) * f + v * f;
      v *= (x - xn[m]) * f;
    }
    if (l) l[j] = v;
    if (dl) dl[j] = d;
  }
}

/* Small dense inverse (Gauss-Jordan, partial pivoting); the reference uses
   LAPACK dgetrf/dgetri (StandAloneFunctions.cpp:144-177). */
static int invert(int n, double *a)
{
  int *piv = (int *) malloc(n * sizeof(int));
  double *b = dalloc((size_t) n * n);
  for (int i = 0; i < n; ++i) b[i * n + i] = 1.0;
  for (int c = 0; c < n; ++c) {
    int p = c;
    for (int r = c + 1; r < n; ++r) if (fabs(a[r * n + c]) > fabs(a[p * n + c])) p = r;
    if (a[p * n + c] == 0.0) { free(piv); free(b); return -1; }
    for (int k = 0; k < n; ++k) {
      double t = a[c * n + k]; a[c * n + k] = a[p * n + k]; a[p * n + k] = t;
      t = b[c * n + k]; b[c * n + k] = b[p * n + k]; b[p * n + k] = t;
    }
    const double d = 1.0 / a[c * n + c];
    for (int k = 0; k < n; ++k) { a[c * n + k] *= d; b[c * n + k] *= d; }
    for (int r = 0; r < n; ++r) {
      if (r == c) continue;
      const double f = a[r * n + c];
      if (f == 0.0) continue;
      for (int k = 0; k < n; ++k) { a[r * n + k] -= f * a[c * n + k]; b[r * n + k] -= f * b[c * n + k]; }
    }
  }
  memcpy(a, b, sizeof(double) * n * n);
  free(piv); free(b);
  return 0;
}

/* ------------------------------------------------------ standard element */
/* StandardElementClass::DataStandardElement (StandardElementClass.cpp:105-175)
   and its helpers :329-911. */
static void standard_element(orc_case *c)
{
  const int K = c->K, M = c->M;
  c->rInt = dalloc(M); c->wInt = dalloc(M); c->rDOF = dalloc(K); c->rGrid = dalloc(c->Kg);
  gauss_legendre(M, c->rInt, c->wInt);
  dof_locations(K, c->prm.dof_location, c->rDOF);
  c->A = dalloc(M * K); c->D = dalloc(M * K); c->AT = dalloc(M * K); c->DT = dalloc(M * K);
  for (int p = 0; p < M; ++p)
    for (int q = 0; q < K; ++q) {
      double v, d; legendre(q, c->rInt[p], &v, &d);
      c->A[p * K + q] = v; c->D[p * K + q] = d;
      c->AT[q * M + p] = v; c->DT[q * M + p] = d;
    }
  c->LD = dalloc(K * K); c->VI = dalloc(K * K); c->DLD = dalloc(K * K);
  for (int p = 0; p < K; ++p)
    for (int q = 0; q < K; ++q) {
      double v, d; legendre(q, c->rDOF[p], &v, &d);
      c->LD[p * K + q] = v; c->VI[p * K + q] = v; c->DLD[p * K + q] = d;
    }
  invert(K, c->VI);
  c->FMin = dalloc(K); c->FMax = dalloc(K); c->dFMin = dalloc(K); c->dFMax = dalloc(K);
  for (int q = 0; q < K; ++q) {
    legendre(q, -1.0, &c->FMin[q], &c->dFMin[q]);
    legendre(q, 1.0, &c->FMax[q], &c->dFMax[q]);
  }
  c->w3 = dalloc(c->nInt); c->w2 = dalloc(c->nInt2D);
  for (int k = 0; k < M; ++k) for (int j = 0; j < M; ++j) {
    c->w2[k * M + j] = c->wInt[j] * c->wInt[k];
    for (int i = 0; i < M; ++i) c->w3[(k * M + j) * M + i] = c->wInt[i] * c->wInt[j] * c->wInt[k];
  }
  c->basisCenter = dalloc(c->nDOFs);
  double *v0 = dalloc(K);
  for (int q = 0; q < K; ++q) { double d; legendre(q, 0.0, &v0[q], &d); }
  for (int k = 0; k < K; ++k) for (int j = 0; j < K; ++j) for (int i = 0; i < K; ++i)
    c->basisCenter[(k * K + j) * K + i] = v0[i] * v0[j] * v0[k];
  free(v0);
  /* Lagrange tables (StandardElementClass::ValuesLagrangeBasis1D,
     StandardElementClass.cpp:327-414) */
  c->Lq = dalloc(M * K); c->dLq = dalloc(M * M);
  for (int p = 0; p < M; ++p) {
    lagrange(K, c->rDOF, c->rInt[p], c->Lq + p * K, NULL);
    lagrange(M, c->rInt, c->rInt[p], NULL, c->dLq + p * M);
  }
  double *dl = dalloc(K);
  lagrange(K, c->rDOF, c->rDOF[0], NULL, dl);
  c->dell0 = -dl[0];
  free(dl);
  /* polynomial-correction matrix (StandardElementClass::
     AssemblePolynomialCorrectionOperators, StandardElementClass.cpp:176-291):
     PC = L (dL^T dL)^-1 dL^T with L the 2D nodal->Gauss interpolation and
     dL = dell0 * L. */
  const int n2 = M * M, k2 = K * K;
  double *L2 = dalloc((size_t) n2 * k2), *LS = dalloc((size_t) k2 * k2), *T = dalloc((size_t) k2 * n2);
  for (int i1 = 0; i1 < M; ++i1) for (int i2 = 0; i2 < M; ++i2)
    for (int j1 = 0; j1 < K; ++j1) for (int j2 = 0; j2 < K; ++j2)
      L2[(i1 * M + i2) * k2 + j1 * K + j2] = c->Lq[i1 * K + j1] * c->Lq[i2 * K + j2];
  const double c0 = c->dell0;
  for (int i = 0; i < k2; ++i)
    for (int j = 0; j < k2; ++j) {
      double t = 0.0;
      for (int q = 0; q < n2; ++q) t += (c0 * L2[q * k2 + i]) * (c0 * L2[q * k2 + j]);
      LS[i * k2 + j] = t;
    }
  invert(k2, LS);
  for (int i = 0; i < k2; ++i)
    for (int j = 0; j < n2; ++j) {
      double t = 0.0;
      for (int q = 0; q < k2; ++q) t += LS[i * k2 + q] * (c0 * L2[j * k2 + q]);
      T[i * n2 + j] = t;
    }
  c->PC = dalloc((size_t) n2 * n2);
  for (int i = 0; i < n2; ++i)
    for (int j = 0; j < n2; ++j) {
      double t = 0.0;
      for (int q = 0; q < k2; ++q) t += L2[i * k2 + q] * T[q * n2 + j];
      c->PC[i * n2 + j] = t;
    }
  free(L2); free(LS); free(T);
}

/* ---------------------------------------------------------------- metrics */
/* ElementClass::LGLDistribution (ElementClass.cpp:3406-3623): classify the
   grid DOF spacing of one element; returns 1 LGL, 0 equidistant, -1 unclear. */
static int lgl_distribution(const orc_case *c, const double *g)
{
  const int Pg = c->prm.n_poly_grid, Kg = Pg + 1;
  double req[16], rlgl[16], ld[16];
  dof_locations(Kg, 1, req); dof_locations(Kg, 2, rlgl);
  int nEq = 0, nLgl = 0, nBad = 0;
  for (int dir = 0; dir < 3; ++dir)
    for (int a = 0; a < Kg; ++a)
      for (int b = 0; b < Kg; ++b) {
        ld[0] = 0.0;
        for (int t = 1; t < Kg; ++t) {
          int i1, i0;
          if (dir == 0) { i1 = (a * Kg + b) * Kg + t; i0 = i1 - 1; }
          else if (dir == 1) { i1 = (a * Kg + t) * Kg + b; i0 = i1 - Kg; }
          else { i1 = (t * Kg + a) * Kg + b; i0 = i1 - Kg * Kg; }
          const double dx = g[i1] - g[i0], dy = g[c->nGrid + i1] - g[c->nGrid + i0],
                       dz = g[2 * c->nGrid + i1] - g[2 * c->nGrid + i0];
          ld[t] = ld[t - 1] + sqrt(dx * dx + dy * dy + dz * dz);
        }
        const double s = 2.0 / ld[Pg];
        double mEq = 0.0, mL = 0.0;
        for (int t = 1; t < Pg; ++t) {
          const double v = -1.0 + s * ld[t];
          mEq = fmax(mEq, fabs(v - req[t])); mL = fmax(mL, fabs(v - rlgl[t]));
        }
        if (mEq <= mL) { ++nEq; if (mEq > 0.05) ++nBad; }
        else { ++nLgl; if (mL > 0.05) ++nBad; }
      }
  if (nEq > nLgl) { if (nEq < 10 * nLgl || nEq < 10 * nBad) return -1; }
  else { if (nLgl < 10 * nEq || nLgl < 10 * nBad) return -1; }
  return nLgl > nEq;
}

/* FinalMetric{I,J,K}Face (ElementClass.cpp:1907-2088): unit normal, area
   factor, inverse metric from dx/dr, dx/ds, dx/dt at one face point. */
static void face_metric(int dir, const double d[9], double out[13])
{
  const double dxdr = d[0], dydr = d[1], dzdr = d[2], dxds = d[3], dyds = d[4], dzds = d[5],
               dxdt = d[6], dydt = d[7], dzdt = d[8];
  double nx, ny, nz;
  if (dir == 0) { nx = dyds * dzdt - dydt * dzds; ny = dxdt * dzds - dxds * dzdt; nz = dxds * dydt - dxdt * dyds; }
  else if (dir == 1) { nx = dydt * dzdr - dydr * dzdt; ny = dxdr * dzdt - dxdt * dzdr; nz = dxdt * dydr - dxdr * dydt; }
  else { nx = dydr * dzds - dyds * dzdr; ny = dxds * dzdr - dxdr * dzds; nz = dxdr * dyds - dxds * dydr; }
  const double len = sqrt(nx * nx + ny * ny + nz * nz);
  out[0] = nx / len; out[1] = ny / len; out[2] = nz / len; out[3] = len;
  const double Jinv = 1.0 / (dxdr * (dyds * dzdt - dzds * dydt) - dxds * (dydr * dzdt - dzdr * dydt)
                           + dxdt * (dydr * dzds - dzdr * dyds));
  out[4] = (dyds * dzdt - dzds * dydt) * Jinv; out[5] = (dzds * dxdt - dxds * dzdt) * Jinv;
  out[6] = (dxds * dydt - dyds * dxdt) * Jinv; out[7] = (dzdr * dydt - dydr * dzdt) * Jinv;
  out[8] = (dxdr * dzdt - dzdr * dxdt) * Jinv; out[9] = (dydr * dxdt - dxdr * dydt) * Jinv;
  out[10] = (dydr * dzds - dzdr * dyds) * Jinv; out[11] = (dzdr * dxds - dxdr * dzds) * Jinv;
  out[12] = (dxdr * dyds - dydr * dxds) * Jinv;
}

/* SolverClass::ComputeMetricTerms (SolverClass.cpp:829-1016) with
   ElementClass::ComputeMetricTerms (ElementClass.cpp:1027-1407) and
   ComputeLengthScales (ElementClass.cpp:927-1022). */
static int metrics(orc_case *c, const double *grid)
{
  const int M = c->M, Kg = c->Kg, nInt = c->nInt, n2 = c->nInt2D, nG = c->nGrid;
  int nLgl = 0, nEq = 0;
  for (int e = 0; e < c->nElem; ++e) {
    const int r = lgl_distribution(c, grid + (size_t) e * 3 * nG);
    if (r < 0) return fail("Not able to determine the point distribution");
    if (r) ++nLgl; else ++nEq;
  }
  if (nLgl && nEq) return fail("Elements with both equidistant and LGL spacing encountered");
  dof_locations(Kg, nLgl ? 2 : 1, c->rGrid);
  /* Lagrange tables of the grid DOFs */
  double *Lg = dalloc(M * Kg), *dLg = dalloc(M * Kg);
  double lf[2][16], dlf[2][16];
  for (int p = 0; p < M; ++p) lagrange(Kg, c->rGrid, c->rInt[p], Lg + p * Kg, dLg + p * Kg);
  for (int q = 0; q < Kg; ++q) { lf[0][q] = (q == 0); lf[1][q] = (q == Kg - 1); }
  lagrange(Kg, c->rGrid, -1.0, NULL, dlf[0]);
  lagrange(Kg, c->rGrid, 1.0, NULL, dlf[1]);

  c->vm = dalloc((size_t) c->nElem * 13 * nInt);
  for (int f = 0; f < 6; ++f) c->sm[f] = dalloc((size_t) c->nElem * 16 * n2);
  c->len = dalloc((size_t) c->nElem * 4);
  c->lenLES = dalloc((size_t) c->nElem);
  c->aveNorm = dalloc((size_t) c->nElem * 9);
  /* metric at the nodal solution DOFs for the SGS model's velocity gradients
     there (ElementClass::ComputeCoorNodalSolDOFs, ElementClass.cpp:820-918) */
  const int K = c->K, nD = c->nDOFs, sgs = c->prm.sgs_model != ORC_SGS_NONE;
  double *LgN = dalloc(K * Kg), *dLgN = dalloc(K * Kg), *dN[3][3];
  for (int p = 0; p < K; ++p) lagrange(Kg, c->rGrid, c->rDOF[p], LgN + p * Kg, dLgN + p * Kg);
  for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) dN[a][b] = dalloc(nD);
  c->vmN = dalloc(sgs ? (size_t) c->nElem * 10 * nD : 1);
  double *xyz[4][3];
  for (int a = 0; a < 4; ++a) for (int b = 0; b < 3; ++b) xyz[a][b] = dalloc(nInt);
  double *w = W(c);
  for (int e = 0; e < c->nElem; ++e) {
    const double *g = grid + (size_t) e * 3 * nG;
    double *vm = c->vm + (size_t) e * 13 * nInt;
    for (int b = 0; b < 3; ++b) {
      tp3(Kg, Kg, Kg, M, M, M, Lg, Lg, Lg, g + b * nG, xyz[0][b], 0, w);
      tp3(Kg, Kg, Kg, M, M, M, dLg, Lg, Lg, g + b * nG, xyz[1][b], 0, w);
      tp3(Kg, Kg, Kg, M, M, M, Lg, dLg, Lg, g + b * nG, xyz[2][b], 0, w);
      tp3(Kg, Kg, Kg, M, M, M, Lg, Lg, dLg, g + b * nG, xyz[3][b], 0, w);
    }
    for (int l = 0; l < nInt; ++l) {
      const double dxdr = xyz[1][0][l], dydr = xyz[1][1][l], dzdr = xyz[1][2][l];
      const double dxds = xyz[2][0][l], dyds = xyz[2][1][l], dzds = xyz[2][2][l];
      const double dxdt = xyz[3][0][l], dydt = xyz[3][1][l], dzdt = xyz[3][2][l];
      const double J = dxdr * (dyds * dzdt - dzds * dydt) - dxds * (dydr * dzdt - dzdr * dydt)
                     + dxdt * (dydr * dzds - dzdr * dyds);
      if (J <= 0.0) return fail("Negative or zero Jacobians encountered");
      const double Ji = 1.0 / J;
      vm[0 * nInt + l] = J;
      vm[1 * nInt + l] = Ji * (dyds * dzdt - dzds * dydt);
      vm[2 * nInt + l] = Ji * (dzds * dxdt - dxds * dzdt);
      vm[3 * nInt + l] = Ji * (dxds * dydt - dyds * dxdt);
      vm[4 * nInt + l] = Ji * (dzdr * dydt - dydr * dzdt);
      vm[5 * nInt + l] = Ji * (dxdr * dzdt - dzdr * dxdt);
      vm[6 * nInt + l] = Ji * (dydr * dxdt - dxdr * dydt);
      vm[7 * nInt + l] = Ji * (dydr * dzds - dzdr * dyds);
      vm[8 * nInt + l] = Ji * (dzdr * dxds - dxdr * dzds);
      vm[9 * nInt + l] = Ji * (dxdr * dyds - dydr * dxds);
      vm[10 * nInt + l] = xyz[0][0][l]; vm[11 * nInt + l] = xyz[0][1][l]; vm[12 * nInt + l] = xyz[0][2][l];
    }
    /* faces: f = 2*dir + side */
    for (int f = 0; f < 6; ++f) {
      const int dir = f / 2, side = f % 2;
      const double *fv = lf[side], *dfv = dlf[side];
      double *sm = c->sm[f] + (size_t) e * 16 * n2;
      double *cc = w + 4 * (size_t) nInt * 3; /* after tp3 scratch */
      double *C[4][3];
      for (int a = 0; a < 4; ++a) for (int b = 0; b < 3; ++b) C[a][b] = cc + (size_t) (a * 3 + b) * n2;
      for (int b = 0; b < 3; ++b) {
        const double *gb = g + b * nG;
        if (dir == 0) {
          tp3(Kg, Kg, Kg, 1, M, M, fv, Lg, Lg, gb, C[0][b], 0, w);
          tp3(Kg, Kg, Kg, 1, M, M, dfv, Lg, Lg, gb, C[1][b], 0, w);
          tp3(Kg, Kg, Kg, 1, M, M, fv, dLg, Lg, gb, C[2][b], 0, w);
          tp3(Kg, Kg, Kg, 1, M, M, fv, Lg, dLg, gb, C[3][b], 0, w);
        } else if (dir == 1) {
          tp3(Kg, Kg, Kg, M, 1, M, Lg, fv, Lg, gb, C[0][b], 0, w);
          tp3(Kg, Kg, Kg, M, 1, M, dLg, fv, Lg, gb, C[1][b], 0, w);
          tp3(Kg, Kg, Kg, M, 1, M, Lg, dfv, Lg, gb, C[2][b], 0, w);
          tp3(Kg, Kg, Kg, M, 1, M, Lg, fv, dLg, gb, C[3][b], 0, w);
        } else {
          tp3(Kg, Kg, Kg, M, M, 1, Lg, Lg, fv, gb, C[0][b], 0, w);
          tp3(Kg, Kg, Kg, M, M, 1, dLg, Lg, fv, gb, C[1][b], 0, w);
          tp3(Kg, Kg, Kg, M, M, 1, Lg, dLg, fv, gb, C[2][b], 0, w);
          tp3(Kg, Kg, Kg, M, M, 1, Lg, Lg, dfv, gb, C[3][b], 0, w);
        }
      }
      for (int l = 0; l < n2; ++l) {
        double d[9], out[13];
        for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) d[a * 3 + b] = C[1 + a][b][l];
        face_metric(dir, d, out);
        for (int q = 0; q < 13; ++q) sm[q * n2 + l] = out[q];
        for (int b = 0; b < 3; ++b) sm[(13 + b) * n2 + l] = C[0][b][l];
      }
    }
    if (sgs) {
      double *vn = c->vmN + (size_t) e * 10 * nD;
      for (int b = 0; b < 3; ++b) {
        tp3(Kg, Kg, Kg, K, K, K, dLgN, LgN, LgN, g + b * nG, dN[0][b], 0, w);
        tp3(Kg, Kg, Kg, K, K, K, LgN, dLgN, LgN, g + b * nG, dN[1][b], 0, w);
        tp3(Kg, Kg, Kg, K, K, K, LgN, LgN, dLgN, g + b * nG, dN[2][b], 0, w);
      }
      for (int l = 0; l < nD; ++l) {
        const double dxdr = dN[0][0][l], dydr = dN[0][1][l], dzdr = dN[0][2][l];
        const double dxds = dN[1][0][l], dyds = dN[1][1][l], dzds = dN[1][2][l];
        const double dxdt = dN[2][0][l], dydt = dN[2][1][l], dzdt = dN[2][2][l];
        const double J = dxdr * (dyds * dzdt - dzds * dydt) - dxds * (dydr * dzdt - dzdr * dydt)
                       + dxdt * (dydr * dzds - dzdr * dyds);
        if (J <= 0.0) return fail("Negative or zero Jacobians encountered");
        const double Ji = 1.0 / J;
        vn[0 * nD + l] = J;
        vn[1 * nD + l] = Ji * (dyds * dzdt - dzds * dydt);
        vn[2 * nD + l] = Ji * (dzds * dxdt - dxds * dzdt);
        vn[3 * nD + l] = Ji * (dxds * dydt - dyds * dxdt);
        vn[4 * nD + l] = Ji * (dzdr * dydt - dydr * dzdt);
        vn[5 * nD + l] = Ji * (dxdr * dzdt - dzdr * dxdt);
        vn[6 * nD + l] = Ji * (dydr * dxdt - dxdr * dydt);
        vn[7 * nD + l] = Ji * (dydr * dzds - dzdr * dyds);
        vn[8 * nD + l] = Ji * (dzdr * dxds - dxdr * dzds);
        vn[9 * nD + l] = Ji * (dxdr * dyds - dydr * dxds);
      }
    }
    /* length scales and averaged normals */
    double vol = 0.0;
    for (int l = 0; l < nInt; ++l) vol += c->w3[l] * vm[l];
    double *ln = c->len + (size_t) e * 4;
    ln[3] = vol;
    /* ElementClass::ComputeLengthScales (ElementClass.cpp:937-940) */
    c->lenLES[e] = pow(vol, 1.0 / 3.0) / c->P;
    for (int dir = 0; dir < 3; ++dir) {
      const double *s0 = c->sm[2 * dir] + (size_t) e * 16 * n2, *s1 = c->sm[2 * dir + 1] + (size_t) e * 16 * n2;
      double an[3] = {0, 0, 0}, area = 0.0;
      for (int l = 0; l < n2; ++l) {
        const double wt = c->w2[l];
        for (int b = 0; b < 3; ++b) an[b] += wt * (s0[b * n2 + l] * s0[3 * n2 + l] + s1[b * n2 + l] * s1[3 * n2 + l]);
        area += wt * (s0[3 * n2 + l] + s1[3 * n2 + l]);
      }
      ln[dir] = 2.0 * vol / area;
      for (int b = 0; b < 3; ++b) c->aveNorm[(size_t) e * 9 + dir * 3 + b] = an[b] / area;
    }
  }
  for (int a = 0; a < 4; ++a) for (int b = 0; b < 3; ++b) free(xyz[a][b]);
  for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) free(dN[a][b]);
  free(Lg); free(dLg); free(LgN); free(dLgN);
  return 0;
}

/* Structured neighbours with periodic wrap; the reference aliases the halo
   ring to the opposite owned element (SolverClass.cpp:1689-1693, :1776-1780,
   :1869-1873, :1956-1960, :2048-2052, :2135-2139). */
static void connectivity(orc_case *c)
{
  c->nbr = (int *) malloc(sizeof(int) * 6 * c->nElem);
  const int n[3] = {c->ni, c->nj, c->nk};
  for (int k = 0; k < c->nk; ++k) for (int j = 0; j < c->nj; ++j) for (int i = 0; i < c->ni; ++i) {
    const int e = i + c->ni * (j + c->nj * k);
    const int idx[3] = {i, j, k};
    for (int d = 0; d < 3; ++d)
      for (int s = 0; s < 2; ++s) {
        int t[3] = {idx[0], idx[1], idx[2]};
        t[d] += s ? 1 : -1;
        int ok = 1;
        if (t[d] < 0 || t[d] >= n[d]) {
          if (c->prm.periodic[d]) t[d] = (t[d] + n[d]) % n[d]; else ok = 0;
        }
        c->nbr[e * 6 + 2 * d + s] = ok ? t[0] + c->ni * (t[1] + c->nj * t[2]) : -1;
      }
  }
}

/* --------------------------------------------------------------- physics */
/* Entropy variables -> rho, u, v, w, p, rE (ElementClass.cpp:4195-4206). */
static void ent2prim(const double *V, double *rho, double *u, double *v, double *w,
                     double *p, double *eKin)
{
  const double gm1 = GAM - 1.0, ov1mg = -1.0 / gm1;
  const double V4Inv = 1.0 / V[4];
  *u = -V4Inv * V[1]; *v = -V4Inv * V[2]; *w = -V4Inv * V[3];
  *eKin = 0.5 * ((*u) * (*u) + (*v) * (*v) + (*w) * (*w));
  const double s = GAM - gm1 * (V[0] - V[4] * (*eKin));
  const double tmp = -V[4] * exp(s);
  *rho = pow(tmp, ov1mg);
  *p = -(*rho) * V4Inv;
}

/* Characteristic-matching-layer source at one integration point
   (ElementClass::VolumeResidual step 2, ElementClass.cpp:3887-4117): wave
   amplitudes of the Cartesian primitive gradients in x, y and z with the
   outgoing eigenvalues of the layer orientation removed, P = S L, mapped to
   conservative variables and weighted with the normalised damping
   functions. V: entropy variables, gx/gy/gz: their Cartesian gradients,
   wgt[3]: -J w sigma_d / sigma_max, nsg[3]: layer orientation signs. */
/* Eddy viscosity mu_SGS = rho nu_t of the subgrid scale model at one point:
   g = (du/dx, dv/dx, dw/dx, du/dy, dv/dy, dw/dy, du/dz, dv/dz, dw/dz), i.e.
   G[i][j] = du_i/dx_j = g[3 j + i]; lenScale the element's LES length scale.
   WALE (CWALEModel::EddyViscosity, SGSModel.cpp:34-113): S the symmetric
   part of G, Sd the traceless symmetric part of G^2,
   nu_t = (Cw len)^2 |Sd|^3 / max(|S|^5 + |Sd|^(5/2), eps), Cw = 0.325, with
   |X|^2 = X_ij X_ij. Vreman (CVremanModel::EddyViscosity, SGSModel.cpp:
   128-210): alpha = G with its trace removed from the diagonal, beta = len^2
   alpha alpha^T, B_beta the sum of beta's principal 2 x 2 minors (clipped at
   0), nu_t = c sqrt(B_beta / max(|alpha|^2, eps)), c = 0.07. */
static double eddy_viscosity(int model, double lenScale, double rho, const double *g)
{
  double G[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) G[i][j] = g[3 * j + i];
  if (model == ORC_SGS_WALE) {
    const double len = 0.325 * lenScale;
    double G2[3][3], tr = 0.0;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) G2[i][j] = G[i][0] * G[0][j] + G[i][1] * G[1][j] + G[i][2] * G[2][j];
    for (int i = 0; i < 3; ++i) tr += G2[i][i];
    double s2 = 0.0, sd2 = 0.0;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        const double S = 0.5 * (G[i][j] + G[j][i]);
        const double Sd = 0.5 * (G2[i][j] + G2[j][i]) - (i == j ? (1.0 / 3.0) * tr : 0.0);
        s2 += S * S; sd2 += Sd * Sd;
      }
    const double sd32 = sd2 * sqrt(sd2);
    const double den = s2 * s2 * sqrt(s2) + sqrt(sd32 * sd2);
    return rho * len * len * sd32 / fmax(den, EPS_SMALL);
  }
  if (model == ORC_SGS_VREMAN) {
    const double tr = (1.0 / 3.0) * (G[0][0] + G[1][1] + G[2][2]);
    double al[3][3], a2 = 0.0;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) { al[i][j] = G[i][j] - (i == j ? tr : 0.0); a2 += al[i][j] * al[i][j]; }
    const double l2 = lenScale * lenScale;
    double be[3][3];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) be[i][j] = l2 * (al[i][0] * al[j][0] + al[i][1] * al[j][1] + al[i][2] * al[j][2]);
    double B = be[0][0] * be[1][1] - be[0][1] * be[0][1] + be[0][0] * be[2][2] - be[0][2] * be[0][2]
             + be[1][1] * be[2][2] - be[1][2] * be[1][2];
    if (B < 0.0) B = 0.0;
    return rho * 0.07 * sqrt(B / fmax(a2, EPS_SMALL));
  }
  return 0.0;
}

static void cml_source(const double *V, const double *gx, const double *gy, const double *gz,
                       const double wgt[3], const double nsg[3], double *out)
{
  const double gm1 = GAM - 1.0, ov1mg = -1.0 / gm1, ovgm1 = -ov1mg;
  const double V4Inv = 1.0 / V[4];
  const double u = -V4Inv * V[1], v = -V4Inv * V[2], w = -V4Inv * V[3];
  const double ek = 0.5 * (u * u + v * v + w * w);
  const double s = GAM - gm1 * (V[0] - V[4] * ek);
  const double rho = pow(-V[4] * exp(s), ov1mg);
  const double p = -rho * V4Inv, rE = -p * ov1mg + rho * ek;
  const double H = rE / rho - V4Inv, a2 = -GAM * V4Inv, a = sqrt(a2);
  const double rhoa = rho * a, ovra = 1.0 / rhoa, ova2 = 1.0 / a2, rova = rho / a;
  const double rhou = rho * u, rhov = rho * v, rhow = rho * w;
  /* entropy -> primitive Jacobian rows */
  const double J11 = rho, J12 = rhou, J13 = rhov, J14 = rhow, J15 = rE;
  const double J22 = -V4Inv, J25 = -V4Inv * u, J33 = -V4Inv, J35 = -V4Inv * v, J44 = -V4Inv, J45 = -V4Inv * w;
  const double J51 = p, J52 = p * u, J53 = p * v, J54 = p * w, J55 = p * H;
  const double *g[3] = {gx, gy, gz};
  double dq[3][5];
  for (int d = 0; d < 3; ++d) {
    const double *G = g[d];
    dq[d][0] = J11 * G[0] + J12 * G[1] + J13 * G[2] + J14 * G[3] + J15 * G[4];
    dq[d][1] = J22 * G[1] + J25 * G[4];
    dq[d][2] = J33 * G[2] + J35 * G[4];
    dq[d][3] = J44 * G[3] + J45 * G[4];
    dq[d][4] = J51 * G[0] + J52 * G[1] + J53 * G[2] + J54 * G[3] + J55 * G[4];
  }
  const double vel[3] = {u, v, w};
  for (int l = 0; l < 5; ++l) out[l] = 0.0;
  for (int d = 0; d < 3; ++d) {
    double l1 = vel[d] - a, l2 = vel[d], l3 = vel[d] + a;
    if (nsg[d] * l1 > 0.0) l1 = 0.0;
    if (nsg[d] * l2 > 0.0) l2 = 0.0;
    if (nsg[d] * l3 > 0.0) l3 = 0.0;
    const double dr = dq[d][0], dp = dq[d][4];
    const double dun = dq[d][1 + d];              /* normal velocity derivative */
    const double La = 0.5 * l1 * (ovra * dp - dun);
    const double Lb = 0.5 * l3 * (ovra * dp + dun);
    const double Le = l2 * (dr - ova2 * dp);      /* entropy wave */
    /* P = S L per direction: continuity, momentum (normal: acoustic, the two
       tangential ones: vorticity), energy (ElementClass.cpp:4038-4062) */
    double P[5];
    P[0] = rova * (Lb + La) + Le;
    P[4] = rhoa * (Lb + La);
    for (int t = 0; t < 3; ++t) P[1 + t] = (t == d) ? (Lb - La) : l2 * dq[d][1 + t];
    const double wgt_d = wgt[d];
    out[0] += wgt_d * P[0];
    out[1] += wgt_d * (u * P[0] + rho * P[1]);
    out[2] += wgt_d * (v * P[0] + rho * P[2]);
    out[3] += wgt_d * (w * P[0] + rho * P[3]);
    out[4] += wgt_d * (ek * P[0] + rhou * P[1] + rhov * P[2] + rhow * P[3] + ovgm1 * P[4]);
  }
}

/* ElementClass::VolumeResidual, steps 1-9 (ElementClass.cpp:3809-4533) for
   entropy variables, with or without an SGS model and a body force. */
static void volume_residual(orc_case *c, int e, int mon, double *mach2max, double *eddymax)
{
  const int K = c->K, M = c->M, nInt = c->nInt, nD = c->nDOFs;
  const double *sol = c->sol + (size_t) e * NVAR * nD;
  const double *vm = c->vm + (size_t) e * 13 * nInt;
  double *res = c->res + (size_t) e * NVAR * nD;
  double *dUdV = c->dUdV + (size_t) e * 15 * nInt;
  double *Wb = W(c) + 8 * (size_t) nInt;
  double *Var = Wb, *Dr = Wb + NVAR * nInt, *Ds = Wb + 2 * NVAR * nInt, *Dt = Wb + 3 * NVAR * nInt;
  double *w = W(c);
  for (int l = 0; l < NVAR; ++l) {
    tp3(K, K, K, M, M, M, c->A, c->A, c->A, sol + l * nD, Var + l * nInt, 0, w);
    tp3(K, K, K, M, M, M, c->D, c->A, c->A, sol + l * nD, Dr + l * nInt, 0, w);
    tp3(K, K, K, M, M, M, c->A, c->D, c->A, sol + l * nD, Ds + l * nInt, 0, w);
    tp3(K, K, K, M, M, M, c->A, c->A, c->D, sol + l * nD, Dt + l * nInt, 0, w);
  }
  /* parametric -> Cartesian gradients; bug_compat uses the metric of
     integration point l (the variable index), ElementClass.cpp:3856-3875 */
  for (int l = 0; l < NVAR; ++l)
    for (int m = 0; m < nInt; ++m) {
      const int q = c->prm.bug_compat ? l : m;
      const double dr = Dr[l * nInt + m], ds = Ds[l * nInt + m], dt = Dt[l * nInt + m];
      Dr[l * nInt + m] = dr * vm[1 * nInt + q] + ds * vm[4 * nInt + q] + dt * vm[7 * nInt + q];
      Ds[l * nInt + m] = dr * vm[2 * nInt + q] + ds * vm[5 * nInt + q] + dt * vm[8 * nInt + q];
      Dt[l * nInt + m] = dr * vm[3 * nInt + q] + ds * vm[6 * nInt + q] + dt * vm[9 * nInt + q];
    }
  for (int i = 0; i < NVAR * nD; ++i) res[i] = 0.0;
  const double gm1 = GAM - 1.0;
  const double factHF = c->mu * GAM / PRANDTL_LAM;
  /* sponge element: CML and damping sources (ElementClass.cpp:3887-4117,
     :4450-4489) */
  const int sp = c->spongeIdx ? c->spongeIdx[e] : -1;
  const int useCml = sp >= 0 && c->prm.cml;
  double *Scml = Wb + 20 * (size_t) nInt, *Ssp = Wb + 25 * (size_t) nInt, *Sbf = Wb + 30 * (size_t) nInt;
  /* non-dimensional body force (ElementClass.cpp:4498-4502) */
  const double pRefInv = 1.0 / c->pRef;
  const double fb[3] = {pRefInv * c->prm.body_force[0], pRefInv * c->prm.body_force[1], pRefInv * c->prm.body_force[2]};
  const double *dfn = sp >= 0 ? c->dampFn + (size_t) sp * 3 * nInt : NULL;
  const double *dst = sp >= 0 ? c->dampState + (size_t) sp * 5 * nInt : NULL;
  double ovs[3] = {0.0, 0.0, 0.0}, nsg[3] = {1.0, 1.0, 1.0};
  if (sp >= 0)
    for (int d = 0; d < 3; ++d) {
      const int fl = c->spongeFlag[sp] >> (2 * d);
      if (fl & 1) nsg[d] = -1.0;
      if (fl & 3) ovs[d] = c->matchScale[(size_t) sp * 3 + d];
    }
  for (int m = 0; m < nInt; ++m) {
    double V[5], rho, u, v, ww, p, eKin;
    for (int l = 0; l < 5; ++l) V[l] = Var[l * nInt + m];
    if (useCml) {
      double gx[5], gy[5], gz[5], o[5];
      for (int l = 0; l < 5; ++l) { gx[l] = Dr[l * nInt + m]; gy[l] = Ds[l * nInt + m]; gz[l] = Dt[l * nInt + m]; }
      const double weight = -vm[m] * c->w3[m];
      const double wgt[3] = {weight * ovs[0] * dfn[m], weight * ovs[1] * dfn[nInt + m], weight * ovs[2] * dfn[2 * nInt + m]};
      cml_source(V, gx, gy, gz, wgt, nsg, o);
      for (int l = 0; l < 5; ++l) Scml[l * nInt + m] = o[l];
    }
    ent2prim(V, &rho, &u, &v, &ww, &p, &eKin);
    const double V4Inv = 1.0 / V[4];
    const double rE = p / gm1 + rho * eKin;
    double gx[5], gy[5], gz[5];
    for (int l = 0; l < 5; ++l) { gx[l] = Dr[l * nInt + m]; gy[l] = Ds[l * nInt + m]; gz[l] = Dt[l * nInt + m]; }
    const double dudx = -V4Inv * (gx[1] + u * gx[4]), dudy = -V4Inv * (gy[1] + u * gy[4]), dudz = -V4Inv * (gz[1] + u * gz[4]);
    const double dvdx = -V4Inv * (gx[2] + v * gx[4]), dvdy = -V4Inv * (gy[2] + v * gy[4]), dvdz = -V4Inv * (gz[2] + v * gz[4]);
    const double dwdx = -V4Inv * (gx[3] + ww * gx[4]), dwdy = -V4Inv * (gy[3] + ww * gy[4]), dwdz = -V4Inv * (gz[3] + ww * gz[4]);
    const double fact = V4Inv * V4Inv / gm1;
    const double dedx = fact * gx[4], dedy = fact * gy[4], dedz = fact * gz[4];
    /* dU/dV (ElementClass.cpp:4228-4253) */
    const double rH = rE + p, ru = rho * u, rv = rho * v, rw = rho * ww;
    double *du = dUdV;
    du[0 * nInt + m] = rho; du[1 * nInt + m] = ru; du[2 * nInt + m] = rv; du[3 * nInt + m] = rw;
    du[4 * nInt + m] = rE; du[5 * nInt + m] = ru * u + p; du[6 * nInt + m] = ru * v; du[7 * nInt + m] = ru * ww;
    du[8 * nInt + m] = rH * u; du[9 * nInt + m] = rv * v + p; du[10 * nInt + m] = rv * ww;
    du[11 * nInt + m] = rH * v; du[12 * nInt + m] = rw * ww + p; du[13 * nInt + m] = rH * ww;
    du[14 * nInt + m] = rE * rH / rho + p * eKin;
    if (mon) {
      const double M2 = rho * (u * u + v * v + ww * ww) / (GAM * p);
      if (M2 > *mach2max) *mach2max = M2;
    }
    if (sp >= 0) {
      const double weight = vm[m] * c->w3[m];
      const double sigma = weight * (dfn[m] + dfn[nInt + m] + dfn[2 * nInt + m]);
      Ssp[m] = sigma * (rho - dst[m]);
      Ssp[nInt + m] = sigma * (rho * u - dst[nInt + m]);
      Ssp[2 * nInt + m] = sigma * (rho * v - dst[2 * nInt + m]);
      Ssp[3 * nInt + m] = sigma * (rho * ww - dst[3 * nInt + m]);
      Ssp[4 * nInt + m] = sigma * (rE - dst[4 * nInt + m]);
    }
    /* eddy viscosity (ElementClass.cpp:4268-4287) with its maximum for the
       monitoring data (:4305-4306); total viscosity and heat-flux factor
       (:4339-4342) */
    double muT = c->mu, fhT = factHF;
    if (c->prm.sgs_model != ORC_SGS_NONE) {
      const double g9[9] = {dudx, dvdx, dwdx, dudy, dvdy, dwdy, dudz, dvdz, dwdz};
      const double ev = eddy_viscosity(c->prm.sgs_model, c->lenLES[e], rho, g9);
      if (mon && ev > *eddymax) *eddymax = ev;
      muT = c->mu + ev;
      fhT = c->mu * (GAM / PRANDTL_LAM) + ev * (GAM / PRANDTL_TURB);
    }
    /* fluxes (ElementClass.cpp:4314-4416) */
    const double weight = -vm[m] * c->w3[m];
    if (c->prm.body_force_set) {
      /* body force source -w J (f, u.f) on the momentum and energy equations
         (ElementClass.cpp:4504-4523) */
      Sbf[m] = weight * fb[0];
      Sbf[nInt + m] = weight * fb[1];
      Sbf[2 * nInt + m] = weight * fb[2];
      Sbf[3 * nInt + m] = weight * (u * fb[0] + v * fb[1] + ww * fb[2]);
    }
    const double divV = LAMBDA_OVER_MU * (dudx + dvdy + dwdz);
    const double txx = muT * (2.0 * dudx + divV), tyy = muT * (2.0 * dvdy + divV), tzz = muT * (2.0 * dwdz + divV);
    const double txy = muT * (dudy + dvdx), txz = muT * (dudz + dwdx), tyz = muT * (dvdz + dwdy);
    const double qx = -fhT * dedx, qy = -fhT * dedy, qz = -fhT * dedz;
    const double fx[5] = {ru, ru * u - txx + p, rv * u - txy, rw * u - txz, (rE + p) * u - (u * txx + v * txy + ww * txz - qx)};
    const double fy[5] = {rv, ru * v - txy, rv * v - tyy + p, rw * v - tyz, (rE + p) * v - (u * txy + v * tyy + ww * tyz - qy)};
    const double fz[5] = {rw, ru * ww - txz, rv * ww - tyz, rw * ww - tzz + p, (rE + p) * ww - (u * txz + v * tyz + ww * tzz - qz)};
    for (int d = 0; d < 3; ++d) {
      const double mx = weight * vm[(1 + 3 * d) * nInt + m], my = weight * vm[(2 + 3 * d) * nInt + m],
                   mz = weight * vm[(3 + 3 * d) * nInt + m];
      double *F = (d == 0) ? Dr : (d == 1) ? Ds : Dt;
      for (int l = 0; l < 5; ++l) F[l * nInt + m] = fx[l] * mx + fy[l] * my + fz[l] * mz;
    }
  }
  if (useCml)
    for (int l = 0; l < NVAR; ++l) tp3(M, M, M, K, K, K, c->AT, c->AT, c->AT, Scml + l * nInt, res + l * nD, 1, w);
  for (int l = 0; l < NVAR; ++l) {
    tp3(M, M, M, K, K, K, c->DT, c->AT, c->AT, Dr + l * nInt, res + l * nD, 1, w);
    tp3(M, M, M, K, K, K, c->AT, c->DT, c->AT, Ds + l * nInt, res + l * nD, 1, w);
    tp3(M, M, M, K, K, K, c->AT, c->AT, c->DT, Dt + l * nInt, res + l * nD, 1, w);
  }
  if (sp >= 0)
    for (int l = 0; l < NVAR; ++l) tp3(M, M, M, K, K, K, c->AT, c->AT, c->AT, Ssp + l * nInt, res + l * nD, 1, w);
  if (c->prm.body_force_set)
    for (int l = 0; l < 4; ++l) tp3(M, M, M, K, K, K, c->AT, c->AT, c->AT, Sbf + l * nInt, res + (l + 1) * nD, 1, w);
}

/* Trace of sol (5 x K^3) and its parametric gradients on face f of an
   element: TensorProductSolAndGrad{I,J,K}Face (include/TensorProductSolAndGrad*Face.hpp). */
static void face_trace(const orc_case *c, int f, const double *sol, double *S, double *Gr, double *Gs, double *Gt)
{
  const int K = c->K, M = c->M, n2 = c->nInt2D, nD = c->nDOFs, dir = f / 2, side = f % 2;
  const double *F = side ? c->FMax : c->FMin, *dF = side ? c->dFMax : c->dFMin;
  double *w = W(c);
  for (int l = 0; l < NVAR; ++l) {
    const double *b = sol + l * nD;
    if (dir == 0) {
      tp3(K, K, K, 1, M, M, F, c->A, c->A, b, S + l * n2, 0, w);
      tp3(K, K, K, 1, M, M, dF, c->A, c->A, b, Gr + l * n2, 0, w);
      tp3(K, K, K, 1, M, M, F, c->D, c->A, b, Gs + l * n2, 0, w);
      tp3(K, K, K, 1, M, M, F, c->A, c->D, b, Gt + l * n2, 0, w);
    } else if (dir == 1) {
      tp3(K, K, K, M, 1, M, c->A, F, c->A, b, S + l * n2, 0, w);
      tp3(K, K, K, M, 1, M, c->D, F, c->A, b, Gr + l * n2, 0, w);
      tp3(K, K, K, M, 1, M, c->A, dF, c->A, b, Gs + l * n2, 0, w);
      tp3(K, K, K, M, 1, M, c->A, F, c->D, b, Gt + l * n2, 0, w);
    } else {
      tp3(K, K, K, M, M, 1, c->A, c->A, F, b, S + l * n2, 0, w);
      tp3(K, K, K, M, M, 1, c->D, c->A, F, b, Gr + l * n2, 0, w);
      tp3(K, K, K, M, M, 1, c->A, c->D, F, b, Gs + l * n2, 0, w);
      tp3(K, K, K, M, M, 1, c->A, c->A, dF, b, Gt + l * n2, 0, w);
    }
  }
}

/* Lift face data into 5 x K^3 residual: TensorProductRes{I,J,K}Face
   (include/TensorProductRes*Face.hpp), normal flux with the face vector and
   the three symmetrising fluxes with the derivative in their slot. */
static void face_lift(const orc_case *c, int f, const double *flux, const double *sx,
                      const double *sy, const double *sz, double *res)
{
  const int K = c->K, M = c->M, n2 = c->nInt2D, nD = c->nDOFs, dir = f / 2, side = f % 2;
  const double *F = side ? c->FMax : c->FMin, *dF = side ? c->dFMax : c->dFMin;
  double *w = W(c);
  const double *AT = c->AT, *DT = c->DT;
  for (int l = 0; l < NVAR; ++l) {
    double *r = res + l * nD;
    if (dir == 0) {
      tp3(1, M, M, K, K, K, F, AT, AT, flux + l * n2, r, 1, w);
      tp3(1, M, M, K, K, K, dF, AT, AT, sx + l * n2, r, 1, w);
      tp3(1, M, M, K, K, K, F, DT, AT, sy + l * n2, r, 1, w);
      tp3(1, M, M, K, K, K, F, AT, DT, sz + l * n2, r, 1, w);
    } else if (dir == 1) {
      tp3(M, 1, M, K, K, K, AT, F, AT, flux + l * n2, r, 1, w);
      tp3(M, 1, M, K, K, K, DT, F, AT, sx + l * n2, r, 1, w);
      tp3(M, 1, M, K, K, K, AT, dF, AT, sy + l * n2, r, 1, w);
      tp3(M, 1, M, K, K, K, AT, F, DT, sz + l * n2, r, 1, w);
    } else {
      tp3(M, M, 1, K, K, K, AT, AT, F, flux + l * n2, r, 1, w);
      tp3(M, M, 1, K, K, K, DT, AT, F, sx + l * n2, r, 1, w);
      tp3(M, M, 1, K, K, K, AT, DT, F, sy + l * n2, r, 1, w);
      tp3(M, M, 1, K, K, K, AT, AT, dF, sz + l * n2, r, 1, w);
    }
  }
}

/* Ismail-Roe entropy-variable flux (FluxesFace.cpp:813-975) at one point;
   n is the unit normal from L to R, area = weight * |n| factor. */
/* InviscidFluxesFace, Ismail-Roe branch for entropy variables
   (FluxesFace.cpp:813-975); factU (factUNorm, 0 on walls) multiplies the
   convective and dissipation terms, not the pressure term. */
static void ismail_roe(const double *VL, const double *VR, const double n[3], double area, double factU,
                       double *flux)
{
  const double beta = 1.0 / 6.0, alphaMax = 0.0;
  const double gam = GAM, gm1 = GAM - 1.0, ovgm1 = 1.0 / gm1;
  const double gp1Ovg = (gam + 1.0) / gam, gm1Ovg = gm1 / gam;
  double rL, uL, vL, wL, pL, eL, rR, uR, vR, wR, pR, eR;
  ent2prim(VL, &rL, &uL, &vL, &wL, &pL, &eL);
  ent2prim(VR, &rR, &uR, &vR, &wR, &pR, &eR);
  const double dV0 = VR[0] - VL[0], dV1 = VR[1] - VL[1], dV2 = VR[2] - VL[2], dV3 = VR[3] - VL[3], dV4 = VR[4] - VL[4];
  const double z0L = sqrt(rL / pL), z0R = sqrt(rR / pR);
  const double z1L = uL * z0L, z1R = uR * z0R, z2L = vL * z0L, z2R = vR * z0R, z3L = wL * z0L, z3R = wR * z0R;
  const double z4L = sqrt(rL * pL), z4R = sqrt(rR * pR);
  const double z0A = 0.5 * (z0L + z0R), z1A = 0.5 * (z1L + z1R), z2A = 0.5 * (z2L + z2R),
               z3A = 0.5 * (z3L + z3R), z4A = 0.5 * (z4L + z4R);
  /* logarithmic means (FluxesFace.cpp:876-897) */
  double zeta = z0L / z0R, f = (zeta - 1.0) / (zeta + 1.0), uu = f * f, F;
  F = (uu < 0.01) ? 1.0 + uu / 3.0 + uu * uu / 5.0 + uu * uu * uu / 7.0 + uu * uu * uu * uu / 9.0 : log(zeta) / (2.0 * f);
  const double z0Log = z0A / F;
  zeta = z4L / z4R; f = (zeta - 1.0) / (zeta + 1.0); uu = f * f;
  F = (uu < 0.01) ? 1.0 + uu / 3.0 + uu * uu / 5.0 + uu * uu * uu / 7.0 + uu * uu * uu * uu / 9.0 : log(zeta) / (2.0 * f);
  const double z4Log = z4A / F;
  const double oz0 = 1.0 / z0A;
  const double rA = z0A * z4Log, p1 = oz0 * z4A, p2 = 0.5 * (gp1Ovg * z4Log / z0Log + gm1Ovg * p1);
  const double uA = oz0 * z1A, vA = oz0 * z2A, wA = oz0 * z3A;
  const double vn = uA * n[0] + vA * n[1] + wA * n[2];
  const double kin = 0.5 * (uA * uA + vA * vA + wA * wA);
  const double a2 = gam * p2 / rA, a = sqrt(a2);
  const double H = a2 * ovgm1 + kin, E = H - p2 / rA;
  const double ova = 1.0 / a, ova2 = 1.0 / a2;
  double abv1 = rA * (uA * dV1 + vA * dV2 + wA * dV3);
  double abv2 = abv1 + rA * (dV0 + H * dV4);
  const double dr = abv1 + rA * (dV0 + E * dV4);
  const double dru = uA * abv2 + p2 * dV1, drv = vA * abv2 + p2 * dV2, drw = wA * abv2 + p2 * dV3;
  const double drE = H * abv1 + rA * E * (dV0 + H * dV4) + p2 * kin * dV4;
  double lam1 = fabs(vn + a), lam2 = fabs(vn - a), lam3 = fabs(vn);
  lam1 *= 0.5 * (1.0 + beta + alphaMax); lam2 *= 0.5 * (1.0 + beta + alphaMax); lam3 *= 0.5;
  abv1 = 0.5 * (lam1 + lam2); abv2 = 0.5 * (lam1 - lam2);
  const double abv3 = abv1 - lam3;
  const double abv4 = gm1 * (kin * dr - uA * dru - vA * drv - wA * drw + drE);
  const double abv5 = n[0] * dru + n[1] * drv + n[2] * drw - vn * dr;
  const double abv6 = abv3 * abv4 * ova2 + abv2 * abv5 * ova;
  const double abv7 = abv2 * abv4 * ova + abv3 * abv5;
  flux[0] = factU * rA * vn - factU * (lam3 * dr + abv6);
  flux[1] = factU * rA * vn * uA + p1 * n[0] - factU * (lam3 * dru + uA * abv6 + n[0] * abv7);
  flux[2] = factU * rA * vn * vA + p1 * n[1] - factU * (lam3 * drv + vA * abv6 + n[1] * abv7);
  flux[3] = factU * rA * vn * wA + p1 * n[2] - factU * (lam3 * drw + wA * abv6 + n[2] * abv7);
  flux[4] = factU * rA * vn * H - factU * (lam3 * drE + H * abv6 + vn * abv7);
  for (int l = 0; l < 5; ++l) flux[l] *= area;
}

/* FluxesInternalFace (FluxesFace.cpp:220-343) with ViscousFluxesFace
   (:1290-1651) and PenAndSymFluxesSymmVar (:1152-1284). Inputs are traces
   with parametric gradients; outputs: flux (normal flux, L->R normal) and the
   parametric symmetrising fluxes for L and R. fn multiplies the unit normal
   of mL (factNorm: +1 interior and max faces, -1 boundary min faces,
   FluxesFace.cpp:105-106, :168-174); a boundary face passes its own metric
   as mL and mR (FluxesFace.cpp:112-147, :191-210). On a wall, factU = 0
   (wallPerm) and, with a prescribed heat flux qWall (hf = 1), the laminar
   heat-flux factor is 0 in the viscous, penalty and symmetrising fluxes and
   w A qWall is added to the energy flux (FluxesFace.cpp:1312-1318,
   :1586-1593). */
static void internal_face_flux(const orc_case *c, double fn, double lenScale, double lenLES,
                               const double *SL, double *GrL, double *GsL, double *GtL, const double *mL,
                               const double *SR, double *GrR, double *GsR, double *GtR, const double *mR,
                               double factU, int hf, double qWall, double *force, double *eddymax,
                               double *flux, double *sxL, double *syL, double *szL,
                               double *sxR, double *syR, double *szR)
{
  const int n2 = c->nInt2D;
  const int bug = c->prm.bug_compat;
  /* Cartesian gradients, averaged (FluxesFace.cpp:263-289) */
  double *gx = W(c) + 64 * (size_t) n2, *gy = gx + 5 * n2, *gz = gy + 5 * n2;
  for (int l = 0; l < 5; ++l)
    for (int m = 0; m < n2; ++m) {
      const int q = bug ? (l < n2 ? l : 0) : m;
      const int i = l * n2 + m;
      const double xL = GrL[i] * mL[4 * n2 + q] + GsL[i] * mL[7 * n2 + q] + GtL[i] * mL[10 * n2 + q];
      const double yL = GrL[i] * mL[5 * n2 + q] + GsL[i] * mL[8 * n2 + q] + GtL[i] * mL[11 * n2 + q];
      const double zL = GrL[i] * mL[6 * n2 + q] + GsL[i] * mL[9 * n2 + q] + GtL[i] * mL[12 * n2 + q];
      const double xR = GrR[i] * mR[4 * n2 + q] + GsR[i] * mR[7 * n2 + q] + GtR[i] * mR[10 * n2 + q];
      const double yR = GrR[i] * mR[5 * n2 + q] + GsR[i] * mR[8 * n2 + q] + GtR[i] * mR[11 * n2 + q];
      const double zR = GrR[i] * mR[6 * n2 + q] + GsR[i] * mR[9 * n2 + q] + GtR[i] * mR[12 * n2 + q];
      gx[i] = 0.5 * (xL + xR); gy[i] = 0.5 * (yL + yR); gz[i] = 0.5 * (zL + zR);
    }
  const double gm1 = GAM - 1.0, ovgm1 = 1.0 / gm1;
  const double mu = c->mu, factHF = hf ? 0.0 : GAM / PRANDTL_LAM, factHFT = hf ? 0.0 : GAM / PRANDTL_TURB;
  const int sgs = c->prm.sgs_model;
  const double tau = (c->P + 1.0) * (c->P + 1.0) * c->prm.rel_penalty / lenScale;
  const double lam = LAMBDA_OVER_MU, lamp1 = lam + 1.0, lamp2 = lam + 2.0;
  for (int m = 0; m < n2; ++m) {
    double VL[5], VR[5], fl[5];
    for (int l = 0; l < 5; ++l) { VL[l] = SL[l * n2 + m]; VR[l] = SR[l * n2 + m]; }
    const double nn[3] = {fn * mL[0 * n2 + m], fn * mL[1 * n2 + m], fn * mL[2 * n2 + m]};
    const double wA = c->w2[m] * mL[3 * n2 + m];
    ismail_roe(VL, VR, nn, wA, factU, fl);
    /* viscous flux on the averaged state */
    double Va[5];
    for (int l = 0; l < 5; ++l) Va[l] = 0.5 * (VL[l] + VR[l]);
    const double V4Inv = 1.0 / Va[4];
    const double u = -V4Inv * Va[1], v = -V4Inv * Va[2], w = -V4Inv * Va[3];
    double dx[5], dy[5], dz[5];
    for (int l = 0; l < 5; ++l) { dx[l] = gx[l * n2 + m]; dy[l] = gy[l * n2 + m]; dz[l] = gz[l * n2 + m]; }
    const double dudx = -V4Inv * (dx[1] + u * dx[4]), dvdx = -V4Inv * (dx[2] + v * dx[4]), dwdx = -V4Inv * (dx[3] + w * dx[4]);
    const double dudy = -V4Inv * (dy[1] + u * dy[4]), dvdy = -V4Inv * (dy[2] + v * dy[4]), dwdy = -V4Inv * (dy[3] + w * dy[4]);
    const double dudz = -V4Inv * (dz[1] + u * dz[4]), dvdz = -V4Inv * (dz[2] + v * dz[4]), dwdz = -V4Inv * (dz[3] + w * dz[4]);
    const double fe = V4Inv * V4Inv * ovgm1;
    const double dedx = fe * dx[4], dedy = fe * dy[4], dedz = fe * dz[4];
    /* eddy viscosity on the density of the averaged entropy variables
       (FluxesFace.cpp:1409-1459, :1481-1494); total viscosity and heat-flux
       factor (:1537-1554); factHF / factHFT are 0 on a prescribed heat flux
       (:1312-1318) */
    double muT = mu, fh = mu * factHF, fhRatio = factHF;
    if (sgs != ORC_SGS_NONE) {
      double rA, uA_, vA_, wA_, pA_, eA_;
      ent2prim(Va, &rA, &uA_, &vA_, &wA_, &pA_, &eA_);
      const double g9[9] = {dudx, dvdx, dwdx, dudy, dvdy, dwdy, dudz, dvdz, dwdz};
      const double ev = eddy_viscosity(sgs, lenLES, rA, g9);
      muT = mu + ev;
      fh = mu * factHF + ev * factHFT;
      fhRatio = fh / muT;
      /* face maximum for the monitoring data (FluxesFace.cpp:207-213, :336-342) */
      if (eddymax && ev > *eddymax) *eddymax = ev;
    }
    const double nx = wA * nn[0], ny = wA * nn[1], nz = wA * nn[2];
    const double lamDiv = LAMBDA_OVER_MU * (dudx + dvdy + dwdz);
    const double txx = muT * (2.0 * dudx + lamDiv), tyy = muT * (2.0 * dvdy + lamDiv), tzz = muT * (2.0 * dwdz + lamDiv);
    const double txy = muT * (dudy + dvdx), txz = muT * (dudz + dwdx), tyz = muT * (dvdz + dwdy);
    const double qx = -fh * dedx, qy = -fh * dedy, qz = -fh * dedz;
    fl[1] -= txx * nx + txy * ny + txz * nz;
    fl[2] -= txy * nx + tyy * ny + tyz * nz;
    fl[3] -= txz * nx + tyz * ny + tzz * nz;
    fl[4] -= (u * txx + v * txy + w * txz - qx) * nx + (u * txy + v * tyy + w * tyz - qy) * ny
           + (u * txz + v * tyz + w * tzz - qz) * nz;
    if (hf) fl[4] += wA * qWall;
    if (force) {
      /* wall forces (FluxesFace.cpp:999-1013, :1590-1612): the averaged
         pressure of the two states and the viscous stress, with the normal
         pointing into the boundary */
      double rL_, uL_, vL_, wL_, pL_, eL_, rR_, uR_, vR_, wR_, pR_, eR_;
      ent2prim(VL, &rL_, &uL_, &vL_, &wL_, &pL_, &eL_);
      ent2prim(VR, &rR_, &uR_, &vR_, &wR_, &pR_, &eR_);
      const double abv = 0.5 * fn * c->w2[m] * mL[3 * n2 + m] * (pL_ + pR_);
      force[0] += abv * mL[0 * n2 + m]; force[1] += abv * mL[1 * n2 + m]; force[2] += abv * mL[2 * n2 + m];
      const double area = fn * c->w2[m] * mL[3 * n2 + m];
      const double ax = area * mL[0 * n2 + m], ay = area * mL[1 * n2 + m], az = area * mL[2 * n2 + m];
      force[3] -= txx * ax + txy * ay + txz * az;
      force[4] -= txy * ax + tyy * ay + tyz * az;
      force[5] -= txz * ax + tyz * ay + tzz * az;
    }
    /* penalty and symmetrising fluxes (FluxesFace.cpp:1152-1284) */
    const double pOvRho = -V4Inv, e = pOvRho * ovgm1;
    const double dV1 = VL[1] - VR[1], dV2 = VL[2] - VR[2], dV3 = VL[3] - VR[3], dV4 = VL[4] - VR[4];
    const double factFlux = pOvRho * c->w2[m] * mL[3 * n2 + m] * muT;
    double sx_ = factFlux * nn[0], sy_ = factFlux * nn[1], sz_ = factFlux * nn[2];
    const double abvE = u * u + v * v + w * w + fhRatio * e;
    const double K11_11 = lamp2, K11_14 = u * lamp2, K11_22 = 1.0, K11_24 = v, K11_33 = 1.0, K11_34 = w, K11_44 = abvE + u * u * lamp1;
    const double K22_11 = 1.0, K22_14 = u, K22_22 = lamp2, K22_24 = v * lamp2, K22_33 = 1.0, K22_34 = w, K22_44 = abvE + v * v * lamp1;
    const double K33_11 = 1.0, K33_14 = u, K33_22 = 1.0, K33_24 = v, K33_33 = lamp2, K33_34 = w * lamp2, K33_44 = abvE + w * w * lamp1;
    const double K12_12 = lam, K12_14 = v * lam, K12_21 = 1.0, K12_24 = u, K12_41 = v, K12_42 = u * lam, K12_44 = u * v * lamp1;
    const double K13_13 = lam, K13_14 = w * lam, K13_31 = 1.0, K13_34 = u, K13_41 = w, K13_43 = u * lam, K13_44 = u * w * lamp1;
    const double K23_23 = lam, K23_24 = w * lam, K23_32 = 1.0, K23_34 = v, K23_42 = w, K23_43 = v * lam, K23_44 = v * w * lamp1;
    const double un = u * nn[0] + v * nn[1] + w * nn[2];
    const double dVn = dV1 * nn[0] + dV2 * nn[1] + dV3 * nn[2];
    const double ttau = tau * factFlux, abvMom = lamp1 * (dVn + un * dV4);
    fl[1] += ttau * (abvMom * nn[0] + dV1 + u * dV4);
    fl[2] += ttau * (abvMom * nn[1] + dV2 + v * dV4);
    fl[3] += ttau * (abvMom * nn[2] + dV3 + w * dV4);
    fl[4] += ttau * (lamp1 * un * dVn + u * dV1 + v * dV2 + w * dV3 + (abvE + lamp1 * un * un) * dV4);
    const double th = 0.5 * c->prm.theta_sym;
    sx_ *= th; sy_ *= th; sz_ *= th;
    double FX[5], FY[5], FZ[5];
    FX[0] = 0.0;
    FX[1] = -(K11_11 * dV1 + K11_14 * dV4) * sx_ - (K12_12 * dV2 + K12_14 * dV4) * sy_ - (K13_13 * dV3 + K13_14 * dV4) * sz_;
    FX[2] = -(K11_22 * dV2 + K11_24 * dV4) * sx_ - (K12_21 * dV1 + K12_24 * dV4) * sy_;
    FX[3] = -(K11_33 * dV3 + K11_34 * dV4) * sx_ - (K13_31 * dV1 + K13_34 * dV4) * sz_;
    FX[4] = -(K11_14 * dV1 + K11_24 * dV2 + K11_34 * dV3 + K11_44 * dV4) * sx_
            - (K12_41 * dV1 + K12_42 * dV2 + K12_44 * dV4) * sy_ - (K13_41 * dV1 + K13_43 * dV3 + K13_44 * dV4) * sz_;
    FY[0] = 0.0;
    FY[1] = -(K12_21 * dV2 + K12_41 * dV4) * sx_ - (K22_11 * dV1 + K22_14 * dV4) * sy_;
    FY[2] = -(K12_12 * dV1 + K12_42 * dV4) * sx_ - (K22_22 * dV2 + K22_24 * dV4) * sy_ - (K23_23 * dV3 + K23_24 * dV4) * sz_;
    FY[3] = -(K22_33 * dV3 + K22_34 * dV4) * sy_ - (K23_32 * dV2 + K23_34 * dV4) * sz_;
    FY[4] = -(K12_14 * dV1 + K12_24 * dV2 + K12_44 * dV4) * sx_
            - (K22_14 * dV1 + K22_24 * dV2 + K22_34 * dV3 + K22_44 * dV4) * sy_ - (K23_42 * dV2 + K23_43 * dV3 + K23_44 * dV4) * sz_;
    FZ[0] = 0.0;
    FZ[1] = -(K13_31 * dV3 + K13_41 * dV4) * sx_ - (K33_11 * dV1 + K33_14 * dV4) * sz_;
    FZ[2] = -(K23_32 * dV3 + K23_42 * dV4) * sy_ - (K33_22 * dV2 + K33_24 * dV4) * sz_;
    FZ[3] = -(K13_13 * dV1 + K13_43 * dV4) * sx_ - (K23_23 * dV2 + K23_43 * dV4) * sy_ - (K33_33 * dV3 + K33_34 * dV4) * sz_;
    FZ[4] = -(K13_14 * dV1 + K13_34 * dV3 + K13_44 * dV4) * sx_
            - (K23_24 * dV2 + K23_34 * dV3 + K23_44 * dV4) * sy_ - (K33_14 * dV1 + K33_24 * dV2 + K33_34 * dV3 + K33_44 * dV4) * sz_;
    for (int l = 0; l < 5; ++l) {
      flux[l * n2 + m] = fl[l];
      const int q = bug ? (l < n2 ? l : 0) : m;
      /* Cartesian symmetrising flux -> parametric, L and R metrics
         (FluxesFace.cpp:316-334) */
      sxL[l * n2 + m] = FX[l] * mL[4 * n2 + q] + FY[l] * mL[5 * n2 + q] + FZ[l] * mL[6 * n2 + q];
      syL[l * n2 + m] = FX[l] * mL[7 * n2 + q] + FY[l] * mL[8 * n2 + q] + FZ[l] * mL[9 * n2 + q];
      szL[l * n2 + m] = FX[l] * mL[10 * n2 + q] + FY[l] * mL[11 * n2 + q] + FZ[l] * mL[12 * n2 + q];
      sxR[l * n2 + m] = FX[l] * mR[4 * n2 + q] + FY[l] * mR[5 * n2 + q] + FZ[l] * mR[6 * n2 + q];
      syR[l * n2 + m] = FX[l] * mR[7 * n2 + q] + FY[l] * mR[8 * n2 + q] + FZ[l] * mR[9 * n2 + q];
      szR[l * n2 + m] = FX[l] * mR[10 * n2 + q] + FY[l] * mR[11 * n2 + q] + FZ[l] * mR[12 * n2 + q];
    }
  }
}

/* Boundary state of a PC-NSCBC face (BCStandardCharacteristicSubfaceClass
   and its outflow / static-inflow / total-inflow ComputeBoundaryState,
   BoundaryState.cpp:996-1291, :1300-1568, :1579-1837, :1848-2241).
   In: the interior trace SL [5][n2] and its parametric gradients
   GL[0..2] (d/dxi, d/deta, d/dzeta, each [5][n2]).
   Out: boundary entropy state SR and its parametric gradients GR[0..2]. */
static void nscbc_state(const orc_case *c, int f, int e, const double *SL, double *const GL[3],
                        double *SR, double *const GR[3])
{
  const int n2 = c->nInt2D, M = c->M, bc = c->prm.bc[f];
  const double fn = (f % 2) ? 1.0 : -1.0;
  const double gm1 = GAM - 1.0, ovgm1 = 1.0 / gm1, govgm1 = GAM * ovgm1;
  const double *sm = c->sm[f] + (size_t) e * 16 * n2;
  const int *ig = c->ig[f];
  const double *dL[3] = {GL[ig[0]], GL[ig[1]], GL[ig[2]]};
  double *dR[3] = {GR[ig[0]], GR[ig[1]], GR[ig[2]]};
  const double *mI[3], *mJ[3], *mK[3];
  for (int q = 0; q < 3; ++q) {
    mI[q] = sm + (size_t) (4 + 3 * ig[0] + q) * n2;
    mJ[q] = sm + (size_t) (4 + 3 * ig[1] + q) * n2;
    mK[q] = sm + (size_t) (4 + 3 * ig[2] + q) * n2;
  }
  /* ConvertEntropyToPrimitive (BoundaryState.cpp:1002-1108): SR <- (rho,u,v,w,p),
     dR[i] <- dQ/dV dL[i] */
  for (int l = 0; l < n2; ++l) {
    const double V0 = SL[l], V1 = SL[n2 + l], V2 = SL[2 * n2 + l], V3 = SL[3 * n2 + l], V4 = SL[4 * n2 + l];
    const double V4Inv = 1.0 / V4;
    const double u = -V4Inv * V1, v = -V4Inv * V2, w = -V4Inv * V3;
    const double ek = 0.5 * (u * u + v * v + w * w);
    const double sEnt = GAM - gm1 * (V0 - V4 * ek);
    const double a2 = -GAM * V4Inv;
    const double rho = pow(-V4 * exp(sEnt), -ovgm1);
    const double p = a2 * rho / GAM;
    const double povrho = -V4Inv;
    const double E = ovgm1 * povrho + ek, H = E + povrho;
    for (int i = 0; i < 3; ++i) {
      const double *dV = dL[i];
      double *dQ = dR[i];
      const double d0 = dV[l], d1 = dV[n2 + l], d2 = dV[2 * n2 + l], d3 = dV[3 * n2 + l], d4 = dV[4 * n2 + l];
      dQ[l] = rho * d0 + rho * u * d1 + rho * v * d2 + rho * w * d3 + rho * E * d4;
      dQ[n2 + l] = povrho * d1 + povrho * u * d4;
      dQ[2 * n2 + l] = povrho * d2 + povrho * v * d4;
      dQ[3 * n2 + l] = povrho * d3 + povrho * w * d4;
      dQ[4 * n2 + l] = p * d0 + p * u * d1 + p * v * d2 + p * w * d3 + p * H * d4;
    }
    SR[l] = rho; SR[n2 + l] = u; SR[2 * n2 + l] = v; SR[3 * n2 + l] = w; SR[4 * n2 + l] = p;
  }
  /* wave amplitudes, incoming ones replaced; reconstructed normal derivative
     in dR[0], B = target - (dQ/dn - Q dell) in dR[1] */
  const double dell = fn * c->dell0;
  const int iPhi = fn > 0.0 ? 0 : 4, iPsi = iPhi == 0 ? 4 : 0;
  const double Mach = c->machAvg[f];
  for (int l = 0; l < n2; ++l) {
    double LL[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    const double rr[3] = {mI[0][l], mI[1][l], mI[2][l]};
    const double magrr = sqrt(rr[0] * rr[0] + rr[1] * rr[1] + rr[2] * rr[2]);
    const double ovmagrr = 1.0 / magrr;
    const double rn[3] = {rr[0] * ovmagrr, rr[1] * ovmagrr, rr[2] * ovmagrr};
    const double rho = SR[l], u = SR[n2 + l], v = SR[2 * n2 + l], w = SR[3 * n2 + l], p = SR[4 * n2 + l];
    const double drhodr = dR[0][l], dudr = dR[0][n2 + l], dvdr = dR[0][2 * n2 + l], dwdr = dR[0][3 * n2 + l],
                 dpdr = dR[0][4 * n2 + l];
    const double ovrho = 1.0 / rho, a2 = GAM * p * ovrho, a = sqrt(a2);
    const double ova2 = 1.0 / a2, rhoa = rho * a, ovrhoa = 1.0 / rhoa, rhova = rho / a;
    const double dsdr = drhodr - ova2 * dpdr;
    const double dw1dr = rn[1] * dwdr - rn[2] * dvdr;
    const double dw2dr = rn[2] * dudr - rn[0] * dwdr;
    const double dw3dr = rn[0] * dvdr - rn[1] * dudr;
    const double durdr = rn[0] * dudr + rn[1] * dvdr + rn[2] * dwdr;
    double lmb1, lmb3;
    if (bc == ORC_BC_OUTFLOW_NSCBC) {
      /* BoundaryState.cpp:1318-1519 */
      const double sign = iPhi == 0 ? -1.0 : 1.0;
      const double Mavg2 = Mach * Mach;
      const double kk = 0.5 * c->prm.outlet_sigma * (1.0 - Mavg2) / c->prm.outlet_len;
      const double bl = c->prm.outlet_beta_l < 0.0 ? Mach : c->prm.outlet_beta_l;
      const double bt = c->prm.outlet_beta_t < 0.0 ? Mach : c->prm.outlet_beta_t;
      const double cbl = c->prm.outlet_one_minus_beta ? fmin(1.0, 1.0 - bl) : fmin(1.0, bl);
      const double cbt = c->prm.outlet_one_minus_beta ? fmin(1.0, 1.0 - bt) : fmin(1.0, bt);
      const double pInf = c->bcData[f][0];
      const double ss[3] = {mJ[0][l], mJ[1][l], mJ[2][l]};
      const double tt[3] = {mK[0][l], mK[1][l], mK[2][l]};
      const double duds = dR[1][n2 + l], dudt = dR[2][n2 + l];
      const double dvds = dR[1][2 * n2 + l], dvdt = dR[2][2 * n2 + l];
      const double dwds = dR[1][3 * n2 + l], dwdt = dR[2][3 * n2 + l];
      const double dpds = dR[1][4 * n2 + l], dpdt = dR[2][4 * n2 + l];
      const double ur = u * rr[0] + v * rr[1] + w * rr[2];
      const double us = u * ss[0] + v * ss[1] + w * ss[2];
      const double ut = u * tt[0] + v * tt[1] + w * tt[2];
      const double arr = a * magrr;
      const double ars = a * (rn[0] * ss[0] + rn[1] * ss[1] + rn[2] * ss[2]);
      const double art = a * (rn[0] * tt[0] + rn[1] * tt[1] + rn[2] * tt[2]);
      lmb1 = ur - arr; lmb3 = ur + arr;
      const double durds = rn[0] * duds + rn[1] * dvds + rn[2] * dwds;
      const double durdt = rn[0] * dudt + rn[1] * dvdt + rn[2] * dwdt;
      LL[0] = 0.5 * lmb1 * (ovrhoa * dpdr - durdr);
      LL[1] = (rn[0] * dsdr - dw1dr);
      LL[2] = (rn[1] * dsdr - dw2dr);
      LL[3] = (rn[2] * dsdr - dw3dr);
      LL[4] = 0.5 * lmb3 * (ovrhoa * dpdr + durdr);
      const double LT = 0.5 * ovrhoa * (us + sign * ars) * dpds + sign * 0.5 * us * durds
                      + 0.5 * ovrhoa * (ut + sign * art) * dpdt + sign * 0.5 * ut * durdt;
      const double TT = -0.5 * a * (ss[0] * duds + ss[1] * dvds + ss[2] * dwds
                                  + tt[0] * dudt + tt[1] * dvdt + tt[2] * dwdt);
      LL[iPhi] = kk * ovrho * magrr * (p - pInf) - cbl * LT + cbt * TT;
    } else if (bc == ORC_BC_INFLOW_STATIC_NSCBC) {
      /* BoundaryState.cpp:1596-1783 */
      const double sign = fn > 0.0 ? -1.0 : 1.0;
      const double Mavg2 = Mach * Mach;
      const double ka = sign * 0.5 * c->prm.inlet_sigma * (1.0 - Mavg2) / c->prm.inlet_len;
      const double ku = c->prm.inlet_sigma / c->prm.inlet_len;
      const double rhoInf = c->bcData[f][0], uInf = c->bcData[f][1], vInf = c->bcData[f][2], wInf = c->bcData[f][3];
      const double ur = u * rr[0] + v * rr[1] + w * rr[2];
      const double arr = a * magrr;
      lmb1 = ur - arr; lmb3 = ur + arr;
      const double lmb2 = ur, ovlmb2 = 1.0 / lmb2;
      if (c->prm.inlet_incoming) {
        LL[0] = 0.5 * lmb1 * (ovrhoa * dpdr - durdr);
        LL[4] = 0.5 * lmb3 * (ovrhoa * dpdr + durdr);
      }
      const double drho = rho - rhoInf, du = u - uInf, dv = v - vInf, dw = w - wInf;
      const double kaol = ku * a * ovlmb2;
      LL[iPhi] = ka * a * (rr[0] * du + rr[1] * dv + rr[2] * dw);
      LL[1] = kaol * (rr[2] * dv - rr[1] * dw + rr[0] * drho);
      LL[2] = kaol * (rr[0] * dw - rr[2] * du + rr[1] * drho);
      LL[3] = kaol * (rr[1] * du - rr[0] * dv + rr[2] * drho);
    } else {
      /* total-condition inflow, BoundaryState.cpp:1866-2184 */
      const double ovrg = 1.0 / RGAS, ovcp = 1.0 / (GAM * RGAS / gm1);
      const double kk = -c->prm.inlet_sigma / c->prm.inlet_len;
      const double PtInf = c->bcData[f][0], TtInf = c->bcData[f][1];
      const double ZtInf = c->bcData[f][3], ZpInf = c->bcData[f][4];
      const double ova = 1.0 / a;
      const double T = p * ovrho * ovrg;
      const double ur = u * rn[0] + v * rn[1] + w * rn[2];
      const double w1 = rn[1] * w - rn[2] * v, w2 = rn[2] * u - rn[0] * w, w3 = rn[0] * v - rn[1] * u;
      lmb1 = ur - a; lmb3 = ur + a;
      const double lmb2 = ur, ovlmb2 = 1.0 / lmb2;
      if (c->prm.inlet_incoming) {
        LL[0] = 0.5 * lmb1 * (ovrhoa * dpdr - durdr);
        LL[4] = 0.5 * lmb3 * (ovrhoa * dpdr + durdr);
      }
      const double ek = 0.5 * (u * u + v * v + w * w);
      const double magU2 = 2.0 * ek, magU = sqrt(magU2);
      const double M2 = magU2 / a2;
      const double ff = 1.0 + 0.5 * gm1 * M2;
      const double Pt = p * pow(ff, govgm1), Tt = T * ff;
      const double Zt = v / magU, Zp = w / magU;
      const double ka = kk * a;
      const double X1 = ka * (Pt - PtInf), X2 = ka * (Tt - TtInf), X3 = ka * (Zt - ZtInf), X4 = ka * (Zp - ZpInf);
      const double PtovTt = Pt / Tt, ovrPtovTt = ovrg * PtovTt, Ttovrho = Tt * ovrho;
      const double govaPt = GAM * ova * Pt, ekovrho = ek * ovrho, gm1ekova = gm1 * ek * ova;
      const double gm1ovaTt = gm1 * ova * Tt, ovmagU2 = 1.0 / magU2;
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
    const double ovlmb1 = 1.0 / lmb1, ovlmb3 = 1.0 / lmb3;
    const double t0 = ovlmb1 * rhova * LL[0] + rn[0] * LL[1] + rn[1] * LL[2] + rn[2] * LL[3] + ovlmb3 * rhova * LL[4];
    const double t1 = -ovlmb1 * rn[0] * LL[0] - rn[2] * LL[2] + rn[1] * LL[3] + ovlmb3 * rn[0] * LL[4];
    const double t2 = -ovlmb1 * rn[1] * LL[0] + rn[2] * LL[1] - rn[0] * LL[3] + ovlmb3 * rn[1] * LL[4];
    const double t3 = -ovlmb1 * rn[2] * LL[0] - rn[1] * LL[1] + rn[0] * LL[2] + ovlmb3 * rn[2] * LL[4];
    const double t4 = ovlmb1 * rhoa * LL[0] + ovlmb3 * rhoa * LL[4];
    dR[0][l] = t0; dR[0][n2 + l] = t1; dR[0][2 * n2 + l] = t2; dR[0][3 * n2 + l] = t3; dR[0][4 * n2 + l] = t4;
    dR[1][l] = t0 - (drhodr - rho * dell);
    dR[1][n2 + l] = t1 - (dudr - u * dell);
    dR[1][2 * n2 + l] = t2 - (dvdr - v * dell);
    dR[1][3 * n2 + l] = t3 - (dwdr - w * dell);
    dR[1][4 * n2 + l] = t4 - (dpdr - p * dell);
  }
  /* ReconstructBoundaryStateAndGradient (BoundaryState.cpp:1114-1291) */
  for (int k = 0; k < 5; ++k)
    for (int i = 0; i < n2; ++i) {
      double t = 0.0;
      for (int j = 0; j < n2; ++j) t += c->PC[i * n2 + j] * dR[1][k * n2 + j];
      SR[k * n2 + i] = t;
    }
  if (fn < 0.0) for (int i = 0; i < 5 * n2; ++i) SR[i] = -SR[i];
  for (int l = 0; l < n2; ++l) {
    const double rho = SR[l], u = SR[n2 + l], v = SR[2 * n2 + l], w = SR[3 * n2 + l], p = SR[4 * n2 + l];
    const double sEnt = log(p / pow(rho, GAM));
    const double dr = dR[0][l], du = dR[0][n2 + l], dv = dR[0][2 * n2 + l], dw = dR[0][3 * n2 + l], dp = dR[0][4 * n2 + l];
    const double ovrho = 1.0 / rho, ovp = 1.0 / p, rovp = rho * ovp;
    const double ek = 0.5 * (u * u + v * v + w * w);
    const double W0 = -rovp * ek + (GAM - sEnt) * ovgm1, W1 = rovp * u, W2 = rovp * v, W3 = rovp * w, W4 = -rovp;
    SR[l] = W0; SR[n2 + l] = W1; SR[2 * n2 + l] = W2; SR[3 * n2 + l] = W3; SR[4 * n2 + l] = W4;
    const double q0 = govgm1 * ovrho - ovp * ek, q4 = ovp * (rovp * ek - ovgm1);
    dR[0][l] = q0 * dr + (-W1) * du + (-W2) * dv + (-W3) * dw + q4 * dp;
    dR[0][n2 + l] = u * ovp * dr + rovp * du + (-ovp * W1) * dp;
    dR[0][2 * n2 + l] = v * ovp * dr + rovp * dv + (-ovp * W2) * dp;
    dR[0][3 * n2 + l] = w * ovp * dr + rovp * dw + (-ovp * W3) * dp;
    dR[0][4 * n2 + l] = -ovp * dr + ovp * rovp * dp;
  }
  /* transverse derivatives of the reconstructed state: along the slow and the
     fast face index; the reference stores the slow one in the first
     transverse slot (bug_compat), the correct slot is the second */
  double *dSlow = c->prm.bug_compat ? dR[1] : dR[2];
  double *dFast = c->prm.bug_compat ? dR[2] : dR[1];
  for (int v = 0; v < 5; ++v)
    for (int i = 0; i < M; ++i)
      for (int j = 0; j < M; ++j) {
        double ds = 0.0, dt = 0.0;
        for (int k = 0; k < M; ++k) {
          ds += c->dLq[i * M + k] * SR[v * n2 + k * M + j];
          dt += c->dLq[j * M + k] * SR[v * n2 + i * M + k];
        }
        dSlow[v * n2 + i * M + j] = ds;
        dFast[v * n2 + i * M + j] = dt;
      }
}

/* Boundary state of a face with a pointwise boundary condition
   (ComputeBoundaryState of BCInflowSubsonic, BCOutflowSubsonic,
   BCHeatFluxWall, BCIsothermalWall, BCInviscidWall for entropy variables,
   BoundaryState.cpp:40-118, :156-251, :252-343, :344-372, :399-604,
   :636-792); m: the face metric (unit normal rows 0-2), fn: factNorm,
   d: the non-dimensional prescribed data of the face. */
static void pointwise_state(const orc_case *c, int bc, const double *m, double fn, const double *d,
                            const double *SL, double *SR)
{
  const int n2 = c->nInt2D;
  const double gm1 = GAM - 1.0, ovgm1 = 1.0 / gm1, govgm1 = GAM * ovgm1, ovg = 1.0 / GAM;
  for (int l = 0; l < n2; ++l) {
    double V[5], R[5];
    for (int q = 0; q < 5; ++q) V[q] = SL[q * n2 + l];
    const double n0 = m[l], n1 = m[n2 + l], n2_ = m[2 * n2 + l];
    if (bc == ORC_BC_WALL_INVISCID) {
      const double Vn = V[1] * n0 + V[2] * n1 + V[3] * n2_;
      R[0] = V[0] - 0.5 * Vn * Vn / V[4];
      R[1] = V[1] - Vn * n0; R[2] = V[2] - Vn * n1; R[3] = V[3] - Vn * n2_;
      R[4] = V[4];
    } else if (bc == ORC_BC_WALL_HEATFLUX || bc == ORC_BC_WALL_ISOTHERMAL) {
      R[0] = V[0] - 0.5 * (V[1] * V[1] + V[2] * V[2] + V[3] * V[3]) / V[4];
      R[1] = R[2] = R[3] = 0.0;
      R[4] = bc == ORC_BC_WALL_HEATFLUX ? V[4] : -c->pRef / (RGAS * c->rhoRef) / d[0];
    } else if (bc == ORC_BC_INFLOW_SUBSONIC) {
      const double V4Inv = 1.0 / V[4];
      const double u = -V4Inv * V[1], v = -V4Inv * V[2], w = -V4Inv * V[3];
      const double a2 = -GAM * V4Inv, a = sqrt(a2);
      const double nx = fn * n0, ny = fn * n1, nz = fn * n2_;
      const double un = u * nx + v * ny + w * nz;
      double riemann = 2.0 * a * ovgm1 + un;
      const double H = a2 * ovgm1 + 0.5 * (u * u + v * v + w * w);
      riemann *= sqrt(d[1] / H);
      const double alpha = nx * d[2] + ny * d[3] + nz * d[4];
      const double aa = 1.0 + 0.5 * gm1 * alpha * alpha;
      const double bb = -gm1 * alpha * riemann;
      const double cc = 0.5 * gm1 * riemann * riemann - 2.0 * d[1];
      double dd = bb * bb - 4.0 * aa * cc; dd = sqrt(dd > 0.0 ? dd : 0.0);
      double qR = (-bb + dd) / (2.0 * aa); qR = qR > 0.0 ? qR : 0.0;
      double qR2 = qR * qR;
      double aR2 = gm1 * (d[1] - 0.5 * qR2);
      double MR2 = qR2 / aR2; MR2 = MR2 < 1.0 ? MR2 : 1.0;
      const double tmp = 1.0 / (1.0 + 0.5 * gm1 * MR2);
      aR2 = gm1 * d[1] * tmp;
      qR2 = MR2 * aR2;
      qR = sqrt(qR2);
      const double pR = d[0] * pow(tmp, govgm1);
      const double rhoR = GAM * pR / aR2;
      const double pRInv = 1.0 / pR;
      const double sR = log(pR / pow(rhoR, GAM));
      R[0] = (GAM - sR) * ovgm1 - 0.5 * pRInv * rhoR * qR2;
      R[1] = rhoR * pRInv * qR * d[2]; R[2] = rhoR * pRInv * qR * d[3]; R[3] = rhoR * pRInv * qR * d[4];
      R[4] = -rhoR * pRInv;
    } else if (bc == ORC_BC_INFLOW_SUPERSONIC || bc == ORC_BC_FARFIELD) {   /* BoundaryState.cpp:605-635, :124-150 */
      for (int q = 0; q < 5; ++q) R[q] = d[q];
    } else if (bc == ORC_BC_OUTFLOW_SUPERSONIC) {  /* BoundaryState.cpp:793-818 */
      for (int q = 0; q < 5; ++q) R[q] = V[q];
    } else {  /* subsonic outflow */
      const double V4Inv = 1.0 / V[4];
      const double u = -V4Inv * V[1], v = -V4Inv * V[2], w = -V4Inv * V[3];
      const double eKin = 0.5 * (u * u + v * v + w * w);
      const double s = GAM - gm1 * (V[0] - V[4] * eKin);
      const double a = sqrt(-GAM * V4Inv);
      const double nx = fn * n0, ny = fn * n1, nz = fn * n2_;
      const double un = u * nx + v * ny + w * nz;
      const double riemann = 2.0 * a * ovgm1 + un;
      const double pR = d[0];
      const double rhoR = pow(pR / exp(s), ovg);
      const double aR = sqrt(GAM * pR / rhoR);
      const double unR = riemann - 2.0 * aR * ovgm1;
      const double uR = u + (unR - un) * nx, vR = v + (unR - un) * ny, wR = w + (unR - un) * nz;
      const double pRInv = 1.0 / pR;
      R[0] = (GAM - s) * ovgm1 - 0.5 * pRInv * rhoR * (uR * uR + vR * vR + wR * wR);
      R[1] = rhoR * pRInv * uR; R[2] = rhoR * pRInv * vR; R[3] = rhoR * pRInv * wR;
      R[4] = -rhoR * pRInv;
    }
    for (int q = 0; q < 5; ++q) SR[q * n2 + l] = R[q];
  }
}

static int is_wall(int bc)
{
  return bc == ORC_BC_WALL_HEATFLUX || bc == ORC_BC_WALL_ISOTHERMAL || bc == ORC_BC_WALL_INVISCID;
}

/* Physical face f of element e: FluxesBoundaryFace (FluxesFace.cpp:16-214)
   and the lift of the boundary branch of ElementClass::{I,J,K}FaceResiduals
   (ElementClass.cpp:2751-2850). An NSCBC face averages the interior and
   reconstructed gradients; a pointwise boundary condition uses the interior
   gradient alone (passed as both sides, FluxesFace.cpp:60-80, :117-147). */
static int boundary_face(orc_case *c, int f, int e, double *buf)
{
  const int n2 = c->nInt2D, nD = c->nDOFs, bc = c->prm.bc[f];
  double *SL = buf, *G0 = SL + 5 * n2, *G1 = G0 + 5 * n2, *G2 = G1 + 5 * n2;
  double *SR = G2 + 5 * n2, *H0 = SR + 5 * n2, *H1 = H0 + 5 * n2, *H2 = H1 + 5 * n2;
  double *flux = H2 + 5 * n2, *sxL = flux + 5 * n2, *syL = sxL + 5 * n2, *szL = syL + 5 * n2;
  double *sxR = szL + 5 * n2, *syR = sxR + 5 * n2, *szR = syR + 5 * n2;
  face_trace(c, f, c->sol + (size_t) e * NVAR * nD, SL, G0, G1, G2);
  const double *m = c->sm[f] + (size_t) e * 16 * n2;
  const double fn = (f % 2) ? 1.0 : -1.0;
  const double lenScale = c->len[(size_t) e * 4 + f / 2], lenLES = c->lenLES[e];
  if (is_nscbc(bc)) {
    double *GL[3] = {G0, G1, G2}, *GR[3] = {H0, H1, H2};
    nscbc_state(c, f, e, SL, GL, SR, GR);
    internal_face_flux(c, fn, lenScale, lenLES, SL, G0, G1, G2, m, SR, H0, H1, H2, m, 1.0, 0, 0.0, NULL,
                       c->monNow ? c->eddyElem + e : NULL, flux, sxL, syL, szL, sxR, syR, szR);
  } else {
    pointwise_state(c, bc, m, fn, c->bcData[f], SL, SR);
    /* forces of the walls when monitoring (SolverClass.cpp:3386, FluxesFace.cpp:98-101) */
    double *force = (c->monNow && c->prm.monitor_forces && is_wall(bc)) ? c->forceElem + (size_t) e * 6 : NULL;
    internal_face_flux(c, fn, lenScale, lenLES, SL, G0, G1, G2, m, SR, G0, G1, G2, m, is_wall(bc) ? 0.0 : 1.0,
                       bc == ORC_BC_WALL_HEATFLUX, c->bcData[f][0], force, c->monNow ? c->eddyElem + e : NULL,
                       flux, sxL, syL, szL, sxR, syR, szR);
  }
  face_lift(c, f, flux, sxL, syL, szL, c->res + (size_t) e * NVAR * nD);
  return 0;
}

/* ElementClass::{I,J,K}FaceResiduals for an interior (incl. periodic) min face
   of element e (ElementClass.cpp:2582-2852 and the J/K analogues). */
static int face_residual(orc_case *c, int dir, int e)
{
  const int n2 = c->nInt2D, nD = c->nDOFs;
  const int L = c->nbr[e * 6 + 2 * dir];
  double *buf = W(c) + 80 * (size_t) n2 + 64;
  if (L < 0) {
    double *rmin = c->resMin[dir] + (size_t) e * NVAR * nD;
    for (int i = 0; i < NVAR * nD; ++i) rmin[i] = 0.0;
    if (boundary_face(c, 2 * dir, e, buf)) return -1;
    if (c->nbr[e * 6 + 2 * dir + 1] < 0 && boundary_face(c, 2 * dir + 1, e, buf)) return -1;
    return 0;
  }
  double *SL = buf, *GrL = SL + 5 * n2, *GsL = GrL + 5 * n2, *GtL = GsL + 5 * n2;
  double *SR = GtL + 5 * n2, *GrR = SR + 5 * n2, *GsR = GrR + 5 * n2, *GtR = GsR + 5 * n2;
  double *flux = GtR + 5 * n2, *sxL = flux + 5 * n2, *syL = sxL + 5 * n2, *szL = syL + 5 * n2;
  double *sxR = szL + 5 * n2, *syR = sxR + 5 * n2, *szR = syR + 5 * n2;
  face_trace(c, 2 * dir, c->sol + (size_t) e * NVAR * nD, SR, GrR, GsR, GtR);
  face_trace(c, 2 * dir + 1, c->sol + (size_t) L * NVAR * nD, SL, GrL, GsL, GtL);
  const double lenScale = fmin(c->len[(size_t) e * 4 + dir], c->len[(size_t) L * 4 + dir]);
  const double lenLES = 0.5 * (c->lenLES[e] + c->lenLES[L]);
  const double *mL = c->sm[2 * dir + 1] + (size_t) L * 16 * n2;
  const double *mR = c->sm[2 * dir] + (size_t) e * 16 * n2;
  internal_face_flux(c, 1.0, lenScale, lenLES, SL, GrL, GsL, GtL, mL, SR, GrR, GsR, GtR, mR, 1.0, 0, 0.0, NULL,
                     c->monNow ? c->eddyElem + e : NULL, flux, sxL, syL, szL, sxR, syR, szR);
  double *rmin = c->resMin[dir] + (size_t) e * NVAR * nD;
  for (int i = 0; i < NVAR * nD; ++i) rmin[i] = 0.0;
  face_lift(c, 2 * dir + 1, flux, sxL, syL, szL, rmin);
  for (int i = 0; i < 5 * n2; ++i) flux[i] = -flux[i];
  face_lift(c, 2 * dir, flux, sxR, syR, szR, c->res + (size_t) e * NVAR * nD);
  if (c->nbr[e * 6 + 2 * dir + 1] < 0 && boundary_face(c, 2 * dir + 1, e, buf)) return -1;
  return 0;
}

/* CG_EntropyVarMassMatrix (CG_MassMatrix.cpp:250-504). */
static int cg_entropy(orc_case *c, int e)
{
  const int K = c->K, M = c->M, nD = c->nDOFs, nInt = c->nInt;
  const double *sol = c->sol + (size_t) e * NVAR * nD;
  const double *Jac = c->vm + (size_t) e * 13 * nInt;
  const double *dUdV = c->dUdV + (size_t) e * 15 * nInt;
  double *rv = c->res + (size_t) e * NVAR * nD;
  double *z = W(c) + 8 * (size_t) nInt, *p = z + 5 * nD, *Ap = p + 5 * nD, *x = Ap + 5 * nD, *pI = x + 5 * nD;
  double *w = W(c);
  const double JacInv = 8.0 / c->len[(size_t) e * 4 + 3];
  double r0 = 0.0;
  for (int i = 0; i < 5 * nD; ++i) { r0 = fmax(r0, fabs(rv[i])); x[i] = 0.0; }
  c->cgIt[e] = 0;
  if (r0 * JacInv > 1e-10) {
    double V[5] = {0, 0, 0, 0, 0};
    for (int l = 0; l < 5; ++l) for (int i = 0; i < nD; ++i) V[l] += c->basisCenter[i] * sol[l * nD + i];
    double rho, u, v, ww, pr, eKin;
    ent2prim(V, &rho, &u, &v, &ww, &pr, &eKin);
    const double gm1 = GAM - 1.0, govgm1 = GAM / gm1, pInv = 1.0 / pr;
    const double a1 = gm1 * rho * pInv * pInv, a2 = eKin * a1;
    double dv[15];
    dv[0] = JacInv * (govgm1 / rho + a2 * eKin); dv[1] = -JacInv * u * a2; dv[2] = -JacInv * v * a2;
    dv[3] = -JacInv * ww * a2; dv[4] = JacInv * (a2 - pInv); dv[5] = JacInv * (a1 * u * u + pInv);
    dv[6] = JacInv * a1 * u * v; dv[7] = JacInv * a1 * u * ww; dv[8] = -JacInv * a1 * u;
    dv[9] = JacInv * (a1 * v * v + pInv); dv[10] = JacInv * a1 * v * ww; dv[11] = -JacInv * a1 * v;
    dv[12] = JacInv * (a1 * ww * ww + pInv); dv[13] = -JacInv * a1 * ww; dv[14] = JacInv * a1;
    static const int S[5][5] = {{0, 1, 2, 3, 4}, {1, 5, 6, 7, 8}, {2, 6, 9, 10, 11}, {3, 7, 10, 12, 13}, {4, 8, 11, 13, 14}};
#define PRECOND()                                                                  \
    for (int i = 0; i < nD; ++i)                                                   \
      for (int a = 0; a < 5; ++a) {                                                \
        double s = 0.0;                                                            \
        for (int b = 0; b < 5; ++b) s += dv[S[a][b]] * rv[b * nD + i];             \
        z[a * nD + i] = s;                                                         \
      }
    PRECOND();
    for (int i = 0; i < 5 * nD; ++i) p[i] = z[i];
    for (int it = 0;; ++it) {
      if (it == 100) return fail("Convergence not reached for CG algorithm");
      double rz = 0.0;
      for (int i = 0; i < 5 * nD; ++i) rz += rv[i] * z[i];
      for (int l = 0; l < 5; ++l) tp3(K, K, K, M, M, M, c->A, c->A, c->A, p + l * nD, pI + l * nInt, 0, w);
      for (int m = 0; m < nInt; ++m) {
        const double wt = c->w3[m] * Jac[m];
        double q[5], o[5];
        for (int l = 0; l < 5; ++l) q[l] = wt * pI[l * nInt + m];
        for (int a = 0; a < 5; ++a) {
          double s = 0.0;
          for (int b = 0; b < 5; ++b) s += dUdV[S[a][b] * nInt + m] * q[b];
          o[a] = s;
        }
        for (int l = 0; l < 5; ++l) pI[l * nInt + m] = o[l];
      }
      for (int l = 0; l < 5; ++l) tp3(M, M, M, K, K, K, c->AT, c->AT, c->AT, pI + l * nInt, Ap + l * nD, 0, w);
      double pAp = 0.0;
      for (int i = 0; i < 5 * nD; ++i) pAp += p[i] * Ap[i];
      const double alpha = rz / pAp;
      double rmax = 0.0;
      for (int i = 0; i < 5 * nD; ++i) { x[i] += alpha * p[i]; rv[i] -= alpha * Ap[i]; rmax = fmax(rmax, fabs(rv[i])); }
      c->cgIt[e] = it + 1;
      if (rmax < 1e-10 * r0) break;
      PRECOND();
      double zAp = 0.0;
      for (int i = 0; i < 5 * nD; ++i) zAp += z[i] * Ap[i];
      const double beta = -alpha * zAp / rz;
      for (int i = 0; i < 5 * nD; ++i) p[i] = beta * p[i] + z[i];
    }
#undef PRECOND
  }
  for (int i = 0; i < 5 * nD; ++i) rv[i] = x[i];
  return 0;
}

/* ------------------------------------------------------- case set-up */
static int is_nscbc(int bc)
{
  return bc == ORC_BC_OUTFLOW_NSCBC || bc == ORC_BC_INFLOW_STATIC_NSCBC || bc == ORC_BC_INFLOW_TOTAL_NSCBC;
}

static int is_physical(int bc) { return bc >= ORC_BC_OUTFLOW_NSCBC && bc <= ORC_BC_FARFIELD; }

static int check_params(const orc_case *c)
{
  const orc_params *p = &c->prm;
  for (int f = 0; f < 6; ++f) {
    const int per = p->periodic[f / 2];
    if (per && p->bc[f] != ORC_BC_PERIODIC) return fail("a periodic direction cannot carry a physical boundary");
    if (!per && !is_physical(p->bc[f])) return fail("unsupported boundary condition on a non-periodic block face");
  }
  if (p->outlet_average_local)
    return fail("\"Outlet use local (element) averaged Mach number\" is not supported");
  if (p->cml && !p->sponge_layer) return fail("a characteristic matching layer needs the sponge layer");
  if (p->sponge_layer) {
    if (p->n_poly_grid != p->n_poly_sol) return fail("sponge layer needs equal grid and solution degree");
    int any = 0;
    for (int f = 0; f < 6; ++f) {
      if (p->sponge_elems[f] < 0) return fail("Sponge layer cannot have negative number of elements");
      any |= p->sponge_elems[f] > 0;
    }
    if (!any) return fail("No sponge layer elements have been detected in a sponge layer simulation.");
  }
  return 0;
}

/* InputParamClass::DetermineNonDimPrimVarFreeStream (InputParamClass.cpp:
   991-1191): non-dimensional (rho, u, v, w, p) of the free stream, from the
   prescribed data when they fix it, else from the free-stream data of the
   parameter file (the direction normalised only when the reference
   conditions came from those data, InputParamClass.cpp:697-707). */
static void free_stream_nondim(const orc_case *c, double q[5])
{
  const orc_params *P = &c->prm;
  const double big = 1e300, gm1 = GAM - 1.0;
  double pr = -1.0, pTot = -1.0, rho = -1.0, T = -1.0, TTot = -1.0;
  double vel[3] = {big, big, big}, dir[3] = {big, big, big};
  for (int f = 0; f < 6; ++f) {
    if (P->periodic[f / 2]) continue;
    const int bc = P->bc[f];
    const double *d = P->bc_data[f];
    if (bc == ORC_BC_OUTFLOW_NSCBC || bc == ORC_BC_OUTFLOW_SUBSONIC) pr = d[0];
    if (bc == ORC_BC_INFLOW_STATIC_NSCBC) { rho = d[0]; vel[0] = d[1]; vel[1] = d[2]; vel[2] = d[3]; }
    if (bc == ORC_BC_INFLOW_TOTAL_NSCBC || bc == ORC_BC_INFLOW_SUBSONIC) {
      pTot = d[0]; TTot = d[1]; dir[0] = d[2]; dir[1] = d[3]; dir[2] = d[4];
    }
    if (bc == ORC_BC_INFLOW_SUPERSONIC) { rho = d[0]; vel[0] = d[1]; vel[1] = d[2]; vel[2] = d[3]; pr = d[4]; }
    if (bc == ORC_BC_WALL_ISOTHERMAL) T = d[0];
  }
  int found = 0;
  if (vel[0] < big && vel[1] < big && vel[2] < big) {
    const int n = (pr > 0.0) + (rho > 0.0) + (T > 0.0);
    if (n >= 2) {
      if (pr <= 0.0) pr = rho * RGAS * T;
      if (rho <= 0.0) rho = pr / (RGAS * T);
      found = 1;
    } else if (n == 1 && T <= 0.0) {
      const double vm2 = vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2];
      if (pr <= 0.0) pr = 0.95 * rho * vm2 / GAM; else rho = GAM * pr / (0.95 * vm2);
      found = 1;
    }
    if (found) { q[0] = rho; q[1] = vel[0]; q[2] = vel[1]; q[3] = vel[2]; q[4] = pr; }
  } else if (dir[0] < big && dir[1] < big && dir[2] < big) {
    const double inv = 1.0 / sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
    double M2 = 0.95 * 0.95;
    if (pTot > 0.0 && pr > 0.0) M2 = (pow(pTot / pr, gm1 / GAM) - 1.0) * 2.0 / gm1;
    else if (TTot > 0.0 && T > 0.0) M2 = (TTot / T - 1.0) * 2.0 / gm1;
    if ((pTot > 0.0) + (TTot > 0.0) >= 2) {
      const double rhoTot = pTot / (RGAS * TTot);
      const double tmp = 1.0 + 0.5 * gm1 * M2;
      pr = pTot / pow(tmp, GAM / gm1);
      rho = rhoTot / pow(tmp, 1.0 / gm1);
      const double vm = sqrt(M2 * GAM * pr / rho);
      q[0] = rho; q[1] = dir[0] * inv * vm; q[2] = dir[1] * inv * vm; q[3] = dir[2] * inv * vm; q[4] = pr;
      found = 1;
    }
  }
  if (!found) {
    double dv[3] = {P->free_stream_dir[0], P->free_stream_dir[1], P->free_stream_dir[2]};
    if (c->refFallback) {
      const double inv = 1.0 / sqrt(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
      for (int i = 0; i < 3; ++i) dv[i] *= inv;
    }
    const double vInf = P->mach * sqrt(GAM * RGAS * P->temp_free_stream);
    const double muDim = c->mu * c->pRef / c->uRef;
    const double rhoInf = P->reynolds * muDim / (P->reynolds_length * vInf);
    q[0] = rhoInf; q[1] = vInf * dv[0]; q[2] = vInf * dv[1]; q[3] = vInf * dv[2];
    q[4] = rhoInf * RGAS * P->temp_free_stream;
  }
  q[0] /= c->rhoRef; q[1] /= c->uRef; q[2] /= c->uRef; q[3] /= c->uRef; q[4] /= c->pRef;
}

/* InputParamClass::ComputeReferenceConditions (InputParamClass.cpp:605-758):
   pressure, total pressure, density, temperature and total temperature are
   taken from the prescribed data of the physical faces, a later face
   overriding an earlier one; and the non-dimensionalisation of the
   prescribed data (ConvertPrescribedData, SubfaceClasses.cpp:739-751,
   :863-915, :1169-1180, :1304-1313, :1385-1400, :1453-1475; an isothermal
   wall keeps its dimensional temperature). */
static void reference_conditions(orc_case *c)
{
  const orc_params *q = &c->prm;
  const double muDim = 1.716e-5;
  double p = -1.0, pTot = -1.0, rho = -1.0, T = -1.0, TTot = -1.0;
  for (int f = 0; f < 6; ++f) {
    if (q->periodic[f / 2]) continue;
    const int bc = q->bc[f];
    if (bc == ORC_BC_OUTFLOW_NSCBC || bc == ORC_BC_OUTFLOW_SUBSONIC) p = q->bc_data[f][0];
    if (bc == ORC_BC_INFLOW_STATIC_NSCBC) rho = q->bc_data[f][0];
    if (bc == ORC_BC_INFLOW_TOTAL_NSCBC || bc == ORC_BC_INFLOW_SUBSONIC) { pTot = q->bc_data[f][0]; TTot = q->bc_data[f][1]; }
    if (bc == ORC_BC_WALL_ISOTHERMAL) T = q->bc_data[f][0];
    if (bc == ORC_BC_INFLOW_SUPERSONIC) { rho = q->bc_data[f][0]; p = q->bc_data[f][4]; }
  }
  c->pRef = -1.0;
  if (p > 0.0) c->pRef = p; else if (pTot > 0.0) c->pRef = pTot;
  c->rhoRef = -1.0;
  if (rho > 0.0) c->rhoRef = rho;
  if (c->rhoRef < 0.0 && c->pRef > 0.0) {
    const double TRef = T > 0.0 ? T : TTot;
    if (TRef > 0.0) c->rhoRef = c->pRef / (RGAS * TRef);
  }
  c->refFallback = 0;
  if (c->pRef < 0.0 || c->rhoRef < 0.0) {
    c->refFallback = 1;
    const double vInf = q->mach * sqrt(GAM * RGAS * q->temp_free_stream);
    c->rhoRef = q->reynolds * muDim / (q->reynolds_length * vInf);
    c->pRef = c->rhoRef * RGAS * q->temp_free_stream;
  }
  c->uRef = sqrt(c->pRef / c->rhoRef);
  c->mu = muDim * c->uRef / c->pRef;
  const double TRef = c->pRef / c->rhoRef;
  const double cp = GAM * RGAS / (GAM - 1.0);
  for (int f = 0; f < 6; ++f) {
    const double *d = q->bc_data[f];
    double *o = c->bcData[f];
    for (int l = 0; l < 5; ++l) o[l] = 0.0;
    if (q->periodic[f / 2]) continue;
    const int bc = q->bc[f];
    if (bc == ORC_BC_OUTFLOW_NSCBC || bc == ORC_BC_OUTFLOW_SUBSONIC) o[0] = d[0] / c->pRef;
    else if (bc == ORC_BC_INFLOW_STATIC_NSCBC) {
      o[0] = d[0] / c->rhoRef; o[1] = d[1] / c->uRef; o[2] = d[2] / c->uRef; o[3] = d[3] / c->uRef;
    } else if (bc == ORC_BC_INFLOW_TOTAL_NSCBC) {
      const double inv = 1.0 / sqrt(d[2] * d[2] + d[3] * d[3] + d[4] * d[4]);
      o[0] = d[0] / c->pRef; o[1] = d[1] / TRef; o[2] = d[2] * inv; o[3] = d[3] * inv; o[4] = d[4] * inv;
    } else if (bc == ORC_BC_INFLOW_SUBSONIC) {
      const double inv = 1.0 / sqrt(d[2] * d[2] + d[3] * d[3] + d[4] * d[4]);
      o[0] = d[0] / c->pRef; o[1] = d[1] * (cp / (c->uRef * c->uRef));
      o[2] = d[2] * inv; o[3] = d[3] * inv; o[4] = d[4] * inv;
    } else if (bc == ORC_BC_WALL_HEATFLUX) o[0] = d[0] / (c->pRef * c->uRef);
    else if (bc == ORC_BC_WALL_ISOTHERMAL) o[0] = d[0];
    else if (bc == ORC_BC_FARFIELD) {
      /* the free stream as entropy variables (SubfaceClasses.cpp:542-622) */
      double fs[5];
      free_stream_nondim(c, fs);
      const double s = log(fs[4] / pow(fs[0], GAM)), pInv = 1.0 / fs[4];
      o[0] = (GAM - s) / (GAM - 1.0) - 0.5 * pInv * fs[0] * (fs[1] * fs[1] + fs[2] * fs[2] + fs[3] * fs[3]);
      o[1] = fs[0] * fs[1] * pInv; o[2] = fs[0] * fs[2] * pInv; o[3] = fs[0] * fs[3] * pInv; o[4] = -fs[0] * pInv;
    } else if (bc == ORC_BC_INFLOW_SUPERSONIC) {
      /* the prescribed state as entropy variables (SubfaceClasses.cpp:1026-1106) */
      const double rho = d[0] / c->rhoRef, u = d[1] / c->uRef, v = d[2] / c->uRef, w = d[3] / c->uRef;
      const double pr = d[4] / c->pRef;
      const double s = log(pr / pow(rho, GAM)), pInv = 1.0 / pr;
      o[0] = (GAM - s) / (GAM - 1.0) - 0.5 * pInv * rho * (u * u + v * v + w * w);
      o[1] = rho * u * pInv; o[2] = rho * v * pInv; o[3] = rho * w * pInv; o[4] = -rho * pInv;
    }
  }
}

/* SolverClass::InitNSCBC (SolverClass.cpp:2447-2497): boundary areas
   (ElementClass::ComputeSurfaceAreaBoundary, ElementClass.cpp:303-327) and
   the face orientation of every NSCBC face (ConfigureParamNSCBC,
   SubfaceClasses.cpp:1325-1349, :1363-1384, with IdentifyBoundaryIndices,
   BoundaryState.cpp:829-925, evaluated at integration point 0). */
static int setup_nscbc(orc_case *c)
{
  const int n2 = c->nInt2D;
  c->nscbc = 0;
  for (int f = 0; f < 6; ++f) {
    c->faceArea[f] = 0.0; c->machAvg[f] = 0.0;
    if (c->prm.periodic[f / 2] || !is_nscbc(c->prm.bc[f])) continue;
    c->nscbc = 1;
    int first = -1;
    for (int e = 0; e < c->nElem; ++e) {
      if (c->nbr[e * 6 + f] >= 0) continue;
      if (first < 0) first = e;
      const double *sm = c->sm[f] + (size_t) e * 16 * n2;
      for (int l = 0; l < n2; ++l) c->faceArea[f] += c->w2[l] * sm[3 * n2 + l];
    }
    const double *sm = c->sm[f] + (size_t) first * 16 * n2;
    int found = -1;
    for (int i = 0; i < 3; ++i) {
      const double *g = sm + (size_t) (4 + 3 * i) * n2;
      const double len = sqrt(g[0] * g[0] + g[n2] * g[n2] + g[2 * n2] * g[2 * n2]);
      const double ct = (sm[0] * g[0] + sm[n2] * g[n2] + sm[2 * n2] * g[2 * n2]) / len;
      if (fabs(ct - 1.0) < 1e-10) found = i;
    }
    if (found < 0) return fail("Could not identify relevant boundary indices!");
    c->ig[f][0] = found;
    c->ig[f][1] = found == 0 ? 1 : 0;
    c->ig[f][2] = found == 2 ? 1 : 2;
  }
  return 0;
}

/* SolverClass::InitSpongeLayer and SetUpSpongeLayer{I,J,K}{MIN,MAX}
   (SpongeLayer.cpp:13-662) with ElementClass::ConfigureSpongeLayer
   (SpongeLayer.cpp:844-935): layer membership by global element index,
   damping function sigma_f * (|x_d - x_0| / width)^alpha_f at the
   integration points, CML scale 1/sigma_f. */
static int setup_sponge(orc_case *c, const double *grid)
{
  if (!c->prm.sponge_layer) return 0;
  const int K = c->K, nInt = c->nInt, nG = c->nGrid;
  const int n[3] = {c->ni, c->nj, c->nk};
  c->spongeIdx = (int *) malloc(sizeof(int) * c->nElem);
  c->nSponge = 0;
  for (int e = 0; e < c->nElem; ++e) {
    const int idx[3] = {e % c->ni, (e / c->ni) % c->nj, e / (c->ni * c->nj)};
    int flag = 0;
    for (int d = 0; d < 3; ++d) {
      if (idx[d] < c->prm.sponge_elems[2 * d]) flag |= 1 << (2 * d);
      if (idx[d] >= n[d] - c->prm.sponge_elems[2 * d + 1]) flag |= 1 << (2 * d + 1);
    }
    c->spongeIdx[e] = flag ? c->nSponge++ : -1;
  }
  c->spongeFlag = (unsigned char *) calloc(c->nSponge ? c->nSponge : 1, 1);
  c->dampFn = dalloc((size_t) c->nSponge * 3 * nInt);
  c->dampState = dalloc((size_t) c->nSponge * 5 * nInt);
  c->matchScale = dalloc((size_t) c->nSponge * 3);
  double *xyz = dalloc((size_t) 3 * nInt);
  for (int e = 0; e < c->nElem; ++e) {
    const int s = c->spongeIdx[e];
    if (s < 0) continue;
    const int idx[3] = {e % c->ni, (e / c->ni) % c->nj, e / (c->ni * c->nj)};
    for (int d = 0; d < 3; ++d) {
      if (idx[d] < c->prm.sponge_elems[2 * d]) c->spongeFlag[s] |= 1 << (2 * d);
      if (idx[d] >= n[d] - c->prm.sponge_elems[2 * d + 1]) c->spongeFlag[s] |= 1 << (2 * d + 1);
    }
  }
  for (int f = 0; f < 6; ++f) {
    if (c->prm.sponge_elems[f] <= 0) continue;
    const int d = f / 2, bit = 1 << f;
    const int stride = d == 0 ? 1 : d == 1 ? K : K * K;
    double lo = 1e300, hi = -1e300;
    for (int e = 0; e < c->nElem; ++e) {
      const int s = c->spongeIdx[e];
      if (s < 0 || !(c->spongeFlag[s] & bit)) continue;
      const double *g = grid + ((size_t) e * 3 + d) * nG;
      lo = fmin(lo, g[0]); hi = fmax(hi, g[c->P * stride]);
    }
    const double ov = 1.0 / (hi - lo), x0 = (f % 2) ? lo : hi;
    const double ss = c->prm.damping_constant[f], aa = c->prm.damping_exponent[f];
    for (int e = 0; e < c->nElem; ++e) {
      const int s = c->spongeIdx[e];
      if (s < 0 || !(c->spongeFlag[s] & bit)) continue;
      tp3(K, K, K, c->M, c->M, c->M, c->Lq, c->Lq, c->Lq, grid + ((size_t) e * 3 + d) * nG, xyz, 0, W(c));
      double *fn = c->dampFn + ((size_t) s * 3 + d) * nInt;
      for (int l = 0; l < nInt; ++l) {
        if ((f % 2) ? xyz[l] < x0 : xyz[l] > x0) { free(xyz); return fail("Inconsistent sponge layer location."); }
        fn[l] = ss * pow(ov * fabs(xyz[l] - x0), aa);
      }
      if (c->prm.cml) c->matchScale[(size_t) s * 3 + d] = fabs(ss) < 1.0e-4 ? 1.0 : 1.0 / ss;
    }
  }
  free(xyz);
  return 0;
}

/* ----------------------------------------------------------------- API */
orc_case *orc_create(const orc_params *prm, const double *elem_grid)
{
  if (prm->n_poly_int_rule < 2 * prm->n_poly_sol) { fail("integration rule not accurate enough"); return NULL; }
  orc_case *c = (orc_case *) calloc(1, sizeof(orc_case));
  c->prm = *prm;
  c->ni = prm->n_elem[0]; c->nj = prm->n_elem[1]; c->nk = prm->n_elem[2];
  c->nElem = c->ni * c->nj * c->nk;
  c->P = prm->n_poly_sol; c->K = c->P + 1; c->M = prm->n_poly_int_rule / 2 + 1;
  c->Kg = prm->n_poly_grid + 1; c->nGrid = c->Kg * c->Kg * c->Kg;
  c->nDOFs = c->K * c->K * c->K; c->nInt = c->M * c->M * c->M; c->nInt2D = c->M * c->M;
  if (check_params(c)) { free(c); return NULL; }
  reference_conditions(c);
  size_t big = (size_t) c->nInt; if ((size_t) c->nGrid > big) big = c->nGrid;
  c->work_size = 200 * big + 400 * (size_t) c->nInt2D + 4096;
  c->nthreads = omp_get_max_threads();
  c->works = (double **) calloc(c->nthreads, sizeof(double *));
  for (int t = 0; t < c->nthreads; ++t) c->works[t] = dalloc(c->work_size);
  standard_element(c);
  if (metrics(c, elem_grid)) { orc_destroy(c); return NULL; }
  connectivity(c);
  if (setup_nscbc(c) || setup_sponge(c, elem_grid)) { orc_destroy(c); return NULL; }
  const size_t ns = (size_t) c->nElem * NVAR * c->nDOFs;
  c->sol = dalloc(ns); c->solOld = dalloc(ns); c->res = dalloc(ns);
  for (int d = 0; d < 3; ++d) c->resMin[d] = dalloc(ns);
  c->forceElem = dalloc((size_t) c->nElem * 6);
  c->eddyElem = dalloc((size_t) c->nElem);
  c->avg = dalloc(c->prm.compute_average ? (size_t) c->nElem * c->nDOFs * 12 : 1);
  c->dUdV = dalloc((size_t) c->nElem * 15 * c->nInt);
  c->cgIt = (int *) calloc(c->nElem, sizeof(int));
  return c;
}

void orc_destroy(orc_case *c)
{
  if (!c) return;
  double *ptrs[] = {c->rInt, c->wInt, c->rDOF, c->rGrid, c->A, c->D, c->AT, c->DT, c->LD, c->VI,
                    c->FMin, c->FMax, c->dFMin, c->dFMax, c->w3, c->w2, c->basisCenter, c->vm,
                    c->sm[0], c->sm[1], c->sm[2], c->sm[3], c->sm[4], c->sm[5], c->len, c->aveNorm,
                    c->sol, c->solOld, c->res, c->resMin[0], c->resMin[1], c->resMin[2], c->dUdV,
                    c->Lq, c->dLq, c->PC, c->dampFn, c->dampState, c->matchScale, c->forceElem, c->avg,
                    c->DLD, c->lenLES, c->vmN, c->eddyElem};
  for (size_t i = 0; i < sizeof ptrs / sizeof ptrs[0]; ++i) free(ptrs[i]);
  if (c->works) for (int t = 0; t < c->nthreads; ++t) free(c->works[t]);
  free(c->works);
  free(c->nbr); free(c->cgIt); free(c->spongeIdx); free(c->spongeFlag); free(c);
}

void orc_sizes(const orc_case *c, int *n_elem, int *n_dofs1d, int *n_int1d)
{
  *n_elem = c->nElem; *n_dofs1d = c->K; *n_int1d = c->M;
}

void orc_reference(const orc_case *c, double out[4])
{
  out[0] = c->uRef; out[1] = c->pRef; out[2] = c->rhoRef; out[3] = c->mu;
}

/* SolverClass::ReadRestartSolution, entropy branch (SolverClass.cpp:3300-3372). */
int orc_set_restart(orc_case *c, const double *nodal)
{
  const int K = c->K, nD = c->nDOFs;
  const double gm1 = GAM - 1.0, ovgm1 = 1.0 / gm1;
  double *node = dalloc(5 * (size_t) nD);
  for (int e = 0; e < c->nElem; ++e) {
    for (int d = 0; d < nD; ++d) {
      const double *s = nodal + ((size_t) e * nD + d) * 5;
      double rho = s[0];
      const double ri = 1.0 / rho;
      double u = ri * s[1], v = ri * s[2], w = ri * s[3];
      double p = gm1 * (s[4] - 0.5 * (u * s[1] + v * s[2] + w * s[3]));
      rho /= c->rhoRef; u /= c->uRef; v /= c->uRef; w /= c->uRef; p /= c->pRef;
      const double sEnt = log(p / pow(rho, GAM)), pInv = 1.0 / p;
      node[0 * nD + d] = (GAM - sEnt) * ovgm1 - 0.5 * rho * (u * u + v * v + w * w) * pInv;
      node[1 * nD + d] = rho * u * pInv; node[2 * nD + d] = rho * v * pInv;
      node[3 * nD + d] = rho * w * pInv; node[4 * nD + d] = -rho * pInv;
    }
    for (int l = 0; l < 5; ++l)
      tp3(K, K, K, K, K, K, c->VI, c->VI, c->VI, node + l * nD, c->sol + ((size_t) e * 5 + l) * nD, 0, W(c));
  }
  free(node);
  c->resValid = 0;
  return 0;
}

void orc_set_modal(orc_case *c, const double *modal)
{
  memcpy(c->sol, modal, sizeof(double) * (size_t) c->nElem * 5 * c->nDOFs);
  c->resValid = 0;
}

void orc_get_modal(const orc_case *c, double *modal)
{
  memcpy(modal, c->sol, sizeof(double) * (size_t) c->nElem * 5 * c->nDOFs);
}

void orc_get_residual(const orc_case *c, double *res)
{
  memcpy(res, c->res, sizeof(double) * (size_t) c->nElem * 5 * c->nDOFs);
}

/* SolverClass::WriteRestartSolution solution branch (SolverClass.cpp:4476-4560). */
void orc_get_nodal(const orc_case *c, double *nodal)
{
  const int K = c->K, nD = c->nDOFs;
  double *node = dalloc(5 * (size_t) nD);
  const double momRef = c->rhoRef * c->uRef;
  for (int e = 0; e < c->nElem; ++e) {
    for (int l = 0; l < 5; ++l)
      tp3(K, K, K, K, K, K, c->LD, c->LD, c->LD, c->sol + ((size_t) e * 5 + l) * nD, node + l * nD, 0, W(c));
    for (int d = 0; d < nD; ++d) {
      double V[5], rho, u, v, w, p, eKin;
      for (int l = 0; l < 5; ++l) V[l] = node[l * nD + d];
      ent2prim(V, &rho, &u, &v, &w, &p, &eKin);
      double *o = nodal + ((size_t) e * nD + d) * 5;
      o[0] = c->rhoRef * rho; o[1] = momRef * rho * u; o[2] = momRef * rho * v; o[3] = momRef * rho * w;
      o[4] = c->pRef * (rho * eKin + p / (GAM - 1.0));
    }
  }
  free(node);
}

/* SolverClass::Residual (SolverClass.cpp:3381-3795). */
int orc_residual(orc_case *c, int mon, double md[8])
{
  double mach2 = 0.0, eddy = 0.0;
  int err = 0;
  if (c->prm.sponge_layer && !c->dampStateSet) return fail("sponge layer damping state not set");
  /* the reference calls AverageDataNSCBC before every Residual
     (VCP3D_CurviLinear.cpp:89, SolverClass.cpp:2405) */
  if (c->nscbc) nscbc_average(c);
  c->monNow = mon;
  if (mon) for (size_t i = 0; i < (size_t) c->nElem * 6; ++i) c->forceElem[i] = 0.0;
  if (mon) for (int e = 0; e < c->nElem; ++e) c->eddyElem[e] = 0.0;
#pragma omp parallel for schedule(static) reduction(max : mach2, eddy)
  for (int e = 0; e < c->nElem; ++e) {
    double m2 = 0.0, ev = 0.0;
    volume_residual(c, e, mon, &m2, &ev);
    mach2 = fmax(mach2, m2);
    eddy = fmax(eddy, ev);
  }
  for (int d = 0; d < 3; ++d) {
#pragma omp parallel for schedule(static) reduction(| : err)
    for (int e = 0; e < c->nElem; ++e)
      if (face_residual(c, d, e)) err |= 1;
  }
  if (err) return -1;
  const size_t nv = (size_t) NVAR * c->nDOFs;
  for (int d = 0; d < 3; ++d)
#pragma omp parallel for schedule(static)
    for (int e = 0; e < c->nElem; ++e) {
      const int R = c->nbr[e * 6 + 2 * d + 1];
      if (R < 0) continue;
      double *r = c->res + e * nv;
      const double *s = c->resMin[d] + R * nv;
      for (size_t i = 0; i < nv; ++i) r[i] += s[i];
    }
#pragma omp parallel for schedule(dynamic, 4) reduction(| : err)
  for (int e = 0; e < c->nElem; ++e)
    if (cg_entropy(c, e)) err |= 1;
  if (err) return -1;
  /* force coefficients: element sums in element order, times
     2 / (gamma Mach^2 area) (SolverClass.cpp:3754-3783) */
  for (int i = 0; i < 6; ++i) md[i] = 0.0;
  if (mon) {
    for (int e = 0; e < c->nElem; ++e)
      for (int i = 0; i < 6; ++i) md[i] += c->forceElem[(size_t) e * 6 + i];
    const double area = c->prm.reference_area > 0.0 ? c->prm.reference_area : 1.0;
    const double fc = 2.0 / (GAM * c->prm.mach * c->prm.mach * area);
    for (int i = 0; i < 6; ++i) md[i] *= fc;
  }
  if (mon) for (int e = 0; e < c->nElem; ++e) eddy = fmax(eddy, c->eddyElem[e]);
  md[6] = mach2; md[7] = mon ? eddy : -1.0;
  c->resValid = 1;
  return 0;
}

/* Dimensional eddy viscosity ev[nDOFs] at the nodal DOFs of element e
   (ElementClass.cpp:1562-1573, :3785-3797): the dimensional nodal primitive
   variables (ComputePrimitiveVariablesNodalDOFs, :1500-1517), the
   dimensional velocity gradients of the entropy variables with the nodal
   metric (ComputeVelocityGradientsNodalDOFs, :1649-1790) and the model with
   the element's LES length scale. node: nodal entropy variables [5][nDOFs];
   gb: scratch of 12 nDOFs. */
static void nodal_eddy(const orc_case *c, int e, const double *node, double *gb, double *ev)
{
  const int K = c->K, nD = c->nDOFs;
  const double *sol = c->sol + (size_t) e * NVAR * nD, *vn = c->vmN + (size_t) e * 10 * nD;
  double *dr = gb, *ds = gb + 4 * nD, *dt = gb + 8 * nD;
  for (int l = 1; l < 5; ++l) {
    tp3(K, K, K, K, K, K, c->DLD, c->LD, c->LD, sol + l * nD, dr + (l - 1) * nD, 0, W(c));
    tp3(K, K, K, K, K, K, c->LD, c->DLD, c->LD, sol + l * nD, ds + (l - 1) * nD, 0, W(c));
    tp3(K, K, K, K, K, K, c->LD, c->LD, c->DLD, sol + l * nD, dt + (l - 1) * nD, 0, W(c));
  }
  for (int d = 0; d < nD; ++d) {
    double V[5], rho, u, v, w, p, eKin;
    for (int l = 0; l < 5; ++l) V[l] = node[l * nD + d];
    ent2prim(V, &rho, &u, &v, &w, &p, &eKin);
    rho *= c->rhoRef; p *= c->pRef;
    const double V4Inv = -p / (c->uRef * c->uRef) / rho;
    const double vel[3] = {u, v, w};
    double gp[3][3];   /* [velocity][r|s|t] */
    for (int i = 0; i < 3; ++i) {
      gp[i][0] = -V4Inv * (dr[i * nD + d] + vel[i] * dr[3 * nD + d]);
      gp[i][1] = -V4Inv * (ds[i * nD + d] + vel[i] * ds[3 * nD + d]);
      gp[i][2] = -V4Inv * (dt[i * nD + d] + vel[i] * dt[3 * nD + d]);
    }
    double g9[9];
    for (int x = 0; x < 3; ++x)
      for (int i = 0; i < 3; ++i)
        g9[3 * x + i] = c->uRef * (gp[i][0] * vn[(1 + x) * nD + d] + gp[i][1] * vn[(4 + x) * nD + d]
                                   + gp[i][2] * vn[(7 + x) * nD + d]);
    ev[d] = eddy_viscosity(c->prm.sgs_model, c->lenLES[e], rho, g9);
  }
}

/* ElementClass::ComputeTimeStep (ElementClass.cpp:1522-1644). */
int orc_time_steps(orc_case *c, double *dt)
{
  static const double f1[] = {2.0, 6.0, 12.0, 20.0, 28.0, 38.0, 50.0, 64.0, 80.0, 98.0};
  static const double f2[] = {6.0, 36.0, 150.0, 420.0, 980.0, 1975.0, 3575.0, 7000.0, 14000.0, 28000.0};
  const int K = c->K, nD = c->nDOFs;
  const double muDim = c->mu * c->pRef / c->uRef;
  const double rvl = fmax(fmax(1.0, 2.0 + LAMBDA_OVER_MU), GAM / PRANDTL_LAM);
  const double rvt = fmax(fmax(1.0, 2.0 + LAMBDA_OVER_MU), GAM / PRANDTL_TURB);
  const int sgs = c->prm.sgs_model != ORC_SGS_NONE;
  int bad = 0;
#pragma omp parallel reduction(| : bad)
  {
    double *node = dalloc(5 * (size_t) nD), *gb = dalloc(12 * (size_t) nD), *ev = dalloc(nD);
#pragma omp for schedule(static)
    for (int e = 0; e < c->nElem; ++e) {
      const double *ln = c->len + (size_t) e * 4, *an = c->aveNorm + (size_t) e * 9;
      const double di = 1.0 / ln[0], dj = 1.0 / ln[1], dk = 1.0 / ln[2];
      const double dl2 = di * di + dj * dj + dk * dk;
      const double radVisOverNu = rvl * dl2, radVisOverNuT = rvt * dl2;
      for (int l = 0; l < 5; ++l)
        tp3(K, K, K, K, K, K, c->LD, c->LD, c->LD, c->sol + ((size_t) e * 5 + l) * nD, node + l * nD, 0, W(c));
      if (sgs) nodal_eddy(c, e, node, gb, ev);
      double radInvMax = 0.0, radVisMax = 0.0, a2Min = 1e10;
      for (int d = 0; d < nD; ++d) {
        double V[5], rho, u, v, w, p, eKin;
        for (int l = 0; l < 5; ++l) V[l] = node[l * nD + d];
        ent2prim(V, &rho, &u, &v, &w, &p, &eKin);
        rho *= c->rhoRef; u *= c->uRef; v *= c->uRef; w *= c->uRef; p *= c->pRef;
        const double ri = 1.0 / rho, a2 = GAM * p * ri;
        const double vI = u * an[0] + v * an[1] + w * an[2];
        const double vJ = u * an[3] + v * an[4] + w * an[5];
        const double vK = u * an[6] + v * an[7] + w * an[8];
        const double a = sqrt(fabs(a2));
        const double ri_ = (fabs(vI) + a) * di + (fabs(vJ) + a) * dj + (fabs(vK) + a) * dk;
        const double rv_ = sgs ? ri * (muDim * radVisOverNu + ev[d] * radVisOverNuT) : ri * (muDim * radVisOverNu);
        radInvMax = fmax(radInvMax, ri_); radVisMax = fmax(radVisMax, rv_); a2Min = fmin(a2Min, a2);
      }
      if (a2Min < 0.0) bad = 1;
      const double dtInv = f1[c->P] * radInvMax + f2[c->P] * radVisMax;
      dt[e] = c->prm.cfl * c->uRef / dtInv;
    }
    free(node); free(gb); free(ev);
  }
  if (bad) return fail("Negative square speed of sound");
  return 0;
}

/* SolverClass::ExecuteTimeSynchronizationStep (SolverClass.cpp:2177-2422). */
int orc_sync_step(orc_case *c, double md[8])
{
  static const double alphaRK[] = {1.0, 3.0 / 4.0, 1.0 / 3.0};
  static const double betaRK[] = {1.0, 1.0 / 4.0, 2.0 / 3.0};
  if (!c->resValid && orc_residual(c, 1, md)) return -1;
  const size_t ns = (size_t) c->nElem * NVAR * c->nDOFs;
  double *dte = dalloc(c->nElem);
  int nSteps = 0, reached = 0;
  double elapsed = 0.0;
  while (!reached) {
    ++c->timeStep; ++nSteps;
    if (orc_time_steps(c, dte)) { free(dte); return -1; }
    double dt = 1e10;
    for (int e = 0; e < c->nElem; ++e) dt = fmin(dt, dte[e]);
    memcpy(c->solOld, c->sol, sizeof(double) * ns);
    const double dtSync = c->prm.dt_synchr * c->uRef;
    if (elapsed + dt >= 0.99 * dtSync) { dt = dtSync - elapsed; reached = 1; }
    elapsed += dt;
    for (int s = 0; s < 3; ++s) {
      const double a = alphaRK[s], b = betaRK[s] * dt;
      for (size_t i = 0; i < ns; ++i) c->sol[i] = (1.0 - a) * c->sol[i] + a * c->solOld[i] - b * c->res[i];
      if (orc_residual(c, reached && s == 2, md)) { free(dte); return -1; }
    }
  }
  free(dte);
  if (c->prm.compute_average) {
    /* SolverClass.cpp:2415-2421 and ElementClass::UpdateAverageData
       (ElementClass.cpp:3741-3782): nodal primitive variables
       (ComputePrimitiveVariablesNodalDOFs, :1500-1517, dimensional) */
    ++c->nAvg;
    const double factNew = 1.0 / c->nAvg, factOld = 1.0 - factNew;
    const int K = c->K, nD = c->nDOFs, sgs = c->prm.sgs_model != ORC_SGS_NONE;
    double *node = dalloc(5 * (size_t) nD), *gb = dalloc(12 * (size_t) nD), *ev = dalloc(nD);
    for (int e = 0; e < c->nElem; ++e) {
      for (int l = 0; l < 5; ++l)
        tp3(K, K, K, K, K, K, c->LD, c->LD, c->LD, c->sol + ((size_t) e * 5 + l) * nD, node + l * nD, 0, W(c));
      /* averaged eddy viscosity (ElementClass.cpp:3785-3803) */
      if (sgs) nodal_eddy(c, e, node, gb, ev);
      for (int d = 0; d < nD; ++d) {
        double V[5], rho, u, v, w, p, eKin;
        for (int l = 0; l < 5; ++l) V[l] = node[l * nD + d];
        ent2prim(V, &rho, &u, &v, &w, &p, &eKin);
        rho *= c->rhoRef; u *= c->uRef; v *= c->uRef; w *= c->uRef; p *= c->pRef;
        const double q[11] = {rho, u, v, w, p, u * u, v * v, w * w, u * v, u * w, v * w};
        double *a = c->avg + ((size_t) e * nD + d) * 12;
        for (int i = 0; i < 11; ++i) a[i] = factOld * a[i] + factNew * q[i];
        if (sgs) a[11] = factOld * a[11] + factNew * ev[d];
      }
    }
    free(node); free(gb); free(ev);
  }
  return nSteps;
}

int orc_get_averages(const orc_case *c, double *avg, int *n_steps)
{
  if (!c->prm.compute_average) return fail("the case was created without compute_average");
  memcpy(avg, c->avg, sizeof(double) * (size_t) c->nElem * c->nDOFs * 12);
  *n_steps = c->nAvg;
  return 0;
}

void orc_cg_iterations(const orc_case *c, int *iters)
{
  memcpy(iters, c->cgIt, sizeof(int) * c->nElem);
}

void orc_get_setup(const orc_case *c, int which, double *out)
{
  const size_t E = c->nElem;
  if (which == 0) memcpy(out, c->vm, sizeof(double) * E * 13 * c->nInt);
  else if (which >= 1 && which <= 6) memcpy(out, c->sm[which - 1], sizeof(double) * E * 16 * c->nInt2D);
  else if (which == 7) memcpy(out, c->len, sizeof(double) * E * 4);
  else if (which == 8) for (size_t i = 0; i < 6 * E; ++i) out[i] = c->nbr[i];
  else if (which == 9 || which == 10) {
    const size_t nv = (which == 9 ? 3 : 5) * (size_t) c->nInt;
    for (size_t e = 0; e < E; ++e) {
      const int s = c->spongeIdx ? c->spongeIdx[e] : -1;
      for (size_t i = 0; i < nv; ++i)
        out[e * nv + i] = s < 0 ? 0.0 : (which == 9 ? c->dampFn : c->dampState)[(size_t) s * nv + i];
    }
  }
}

void orc_set_setup(orc_case *c, int which, const double *in)
{
  const size_t E = c->nElem;
  if (which == 0) memcpy(c->vm, in, sizeof(double) * E * 13 * c->nInt);
  else if (which >= 1 && which <= 6) memcpy(c->sm[which - 1], in, sizeof(double) * E * 16 * c->nInt2D);
  else if (which == 7) {
    memcpy(c->len, in, sizeof(double) * E * 4);
    for (size_t e = 0; e < E; ++e) c->lenLES[e] = pow(c->len[e * 4 + 3], 1.0 / 3.0) / c->P;
  }
}

double orc_eddy_viscosity(int model, double len, double rho, const double g[9])
{
  return eddy_viscosity(model, len, rho, g);
}

void orc_get_table(const orc_case *c, int which, double *out)
{
  const int K = c->K, M = c->M;
  switch (which) {
    case 0: memcpy(out, c->rInt, sizeof(double) * M); break;
    case 1: memcpy(out, c->wInt, sizeof(double) * M); break;
    case 2: for (int q = 0; q < K; ++q) for (int p = 0; p < M; ++p) out[q * M + p] = c->A[p * K + q]; break;
    case 3: for (int q = 0; q < K; ++q) for (int p = 0; p < M; ++p) out[q * M + p] = c->D[p * K + q]; break;
    case 4: memcpy(out, c->rDOF, sizeof(double) * K); break;
    case 5: memcpy(out, c->rGrid, sizeof(double) * c->Kg); break;
    case 6: memcpy(out, c->PC, sizeof(double) * M * M * M * M); break;
    case 7: memcpy(out, c->dLq, sizeof(double) * M * M); break;
    case 8: out[0] = c->dell0; break;
  }
}

/* SolverClass::ReadAveragedSolutionSpongeLayer + ElementClass::
   InitializeDampingState (SpongeLayer.cpp:665-838, :940-997): averaged
   nodal primitive state interpolated to the integration points, made
   non-dimensional and converted to conservative variables. */
int orc_set_damping_state(orc_case *c, const double *nodal_prim)
{
  if (!c->prm.sponge_layer) return fail("no sponge layer in this case");
  const int K = c->K, M = c->M, nD = c->nDOFs, nInt = c->nInt;
  double *node = dalloc(5 * (size_t) nD), *q = dalloc(5 * (size_t) nInt);
  const double ovgm1 = 1.0 / (GAM - 1.0);
  for (int e = 0; e < c->nElem; ++e) {
    const int s = c->spongeIdx[e];
    if (s < 0) continue;
    for (int d = 0; d < nD; ++d)
      for (int l = 0; l < 5; ++l) node[l * nD + d] = nodal_prim[((size_t) e * nD + d) * 5 + l];
    for (int l = 0; l < 5; ++l) tp3(K, K, K, M, M, M, c->Lq, c->Lq, c->Lq, node + l * nD, q + l * nInt, 0, W(c));
    double *ds = c->dampState + (size_t) s * 5 * nInt;
    for (int m = 0; m < nInt; ++m) {
      const double rho = q[m] / c->rhoRef, u = q[nInt + m] / c->uRef, v = q[2 * nInt + m] / c->uRef,
                   w = q[3 * nInt + m] / c->uRef, p = q[4 * nInt + m] / c->pRef;
      ds[m] = rho; ds[nInt + m] = rho * u; ds[2 * nInt + m] = rho * v; ds[3 * nInt + m] = rho * w;
      ds[4 * nInt + m] = p * ovgm1 + 0.5 * rho * (u * u + v * v + w * w);
    }
  }
  free(node); free(q);
  c->dampStateSet = 1;
  c->resValid = 0;
  return 0;
}

/* SolverClass::AverageDataNSCBC (SolverClass.cpp:2502-2581) with
   ElementClass::AverageMachNSCBC (ElementClass.cpp:365-479) and
   WeightedMachElement (BoundaryState.cpp:935-994): area-weighted mean of
   the face-normal Mach number u.grad(xi_n)/(|grad(xi_n)| a) over each NSCBC
   block face. */
static void nscbc_average(orc_case *c)
{
  const int n2 = c->nInt2D, nD = c->nDOFs;
  for (int f = 0; f < 6; ++f) {
    double sum = 0.0;
    if (!c->prm.periodic[f / 2] && is_nscbc(c->prm.bc[f])) {
#pragma omp parallel reduction(+ : sum)
      {
        double *S = dalloc(5 * (size_t) n2), *G = dalloc(15 * (size_t) n2);
#pragma omp for schedule(static)
        for (int e = 0; e < c->nElem; ++e) {
          if (c->nbr[e * 6 + f] >= 0) continue;
          face_trace(c, f, c->sol + (size_t) e * NVAR * nD, S, G, G + 5 * n2, G + 10 * n2);
          const double *sm = c->sm[f] + (size_t) e * 16 * n2;
          const double *rr = sm + (size_t) (4 + 3 * c->ig[f][0]) * n2;
          for (int l = 0; l < n2; ++l) {
            const double rx = rr[l], ry = rr[n2 + l], rz = rr[2 * n2 + l];
            const double mag = sqrt(rx * rx + ry * ry + rz * rz);
            const double V4Inv = 1.0 / S[4 * n2 + l];
            const double u = -V4Inv * S[n2 + l], v = -V4Inv * S[2 * n2 + l], w = -V4Inv * S[3 * n2 + l];
            const double a = sqrt(-GAM * V4Inv);
            const double urr = (u * rx + v * ry + w * rz) / mag;
            sum += c->w2[l] * sm[3 * n2 + l] * (urr / a);
          }
        }
        free(S); free(G);
      }
    }
    c->machAvg[f] = fabs(c->faceArea[f]) > 1e-30 ? sum / c->faceArea[f] : sum;
  }
}

void orc_nscbc_average(orc_case *c, double mach_avg[6])
{
  nscbc_average(c);
  for (int f = 0; f < 6; ++f) mach_avg[f] = c->machAvg[f];
}

/* ------------------------------------------------------------ test hooks */
void orc_tensor_apply(int nix, int niy, int niz, int nox, int noy, int noz,
                      const double *ox, const double *oy, const double *oz,
                      const double *in, double *out)
{
  size_t n = (size_t) niz * niy * nox + (size_t) niz * noy * nox + 16;
  double *w = dalloc(n);
  tp3(nix, niy, niz, nox, noy, noz, ox, oy, oz, in, out, 0, w);
  free(w);
}

/* Solves the element's entropy mass-matrix system for rhs [5][nDOFs] with
   the flexible PCG of CG_EntropyVarMassMatrix; dU/dV is the one stored by
   the last orc_residual call. Returns the iteration count or <0. */
int orc_cg_solve(orc_case *c, int e, const double *rhs, double *x)
{
  const size_t nv = (size_t) NVAR * c->nDOFs;
  double *r = c->res + (size_t) e * nv;
  double *save = dalloc(nv);
  memcpy(save, r, sizeof(double) * nv);
  memcpy(r, rhs, sizeof(double) * nv);
  const int rc = cg_entropy(c, e);
  memcpy(x, r, sizeof(double) * nv);
  memcpy(r, save, sizeof(double) * nv);
  free(save);
  return rc < 0 ? rc : c->cgIt[e];
}

/* dU/dV at the integration points of element e, [15][nInt]. */
void orc_get_dudv(const orc_case *c, int e, double *out)
{
  memcpy(out, c->dUdV + (size_t) e * 15 * c->nInt, sizeof(double) * 15 * c->nInt);
}
