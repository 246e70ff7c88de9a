// Device-side metric terms of a block (geometry.hip).
#pragma once

#include <hip/hip_runtime.h>

namespace vcp {

// Inputs and outputs of the geometry kernel. tab: Lg [M][Kg] and dLg [M][Kg]
// (grid-DOF Lagrange basis and derivative at the Gauss points), lf [2][Kg]
// and dlf [2][Kg] (basis and derivative at r = -1, +1), w2 [M^2], w3 [M^3]
// (the setup tables, device copies). Outputs in the residual kernels'
// layouts: vgeo [nElem][kVolGeo][M^3], fgeo [nElem][3][kFaceGeo][M^2],
// econst [nElem][kElemConst] (LES length scale left 0); flag bit 0 is set for
// a non-positive Jacobian.
struct GeoArgs {
  const double *grid;     // [nElem][3][Kg^3]
  const double *tab;
  const int *nbr;         // [nElem][6]
  double *vgeo, *fgeo, *econst;
  int *flag;
  int nElem, Kg;
};

hipError_t launch_geometry(const GeoArgs &a, int M, hipStream_t s);

}  // namespace vcp
