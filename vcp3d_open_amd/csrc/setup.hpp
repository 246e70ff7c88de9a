// Host-side setup of a VCP3D block for the GPU path: 1D operator tables,
// metric terms at volume and face integration points, neighbour maps and
// reference conditions. Runs once per case, in FP64, on the host.
#pragma once

#include <string>
#include <vector>

namespace vcp {

constexpr int kNVar = 5;
constexpr double kGamma = 1.4;            // VCP3D_CurviLinear.hpp:174
constexpr double kPrandtlLam = 0.72;      // VCP3D_CurviLinear.hpp:175
constexpr double kPrandtlTurb = 0.90;     // VCP3D_CurviLinear.hpp:176
constexpr double kRGas = 287.058;         // VCP3D_CurviLinear.hpp:177
constexpr double kLambdaOverMu = -2.0 / 3.0;
constexpr double kMuDim = 1.716e-5;       // InputParamClass.cpp:607

// Physical block-face kinds (one subface per block face).
enum : int { kBCPeriodic = 0, kBCOutflowNSCBC = 1, kBCInflowStaticNSCBC = 2, kBCInflowTotalNSCBC = 3,
              kBCInflowSubsonic = 4, kBCOutflowSubsonic = 5, kBCWallHeatFlux = 6, kBCWallIsothermal = 7,
              kBCWallInviscid = 8, kBCInflowSupersonic = 9, kBCOutflowSupersonic = 10, kBCFarfield = 11 };

inline bool bc_is_nscbc(int bc) { return bc >= kBCOutflowNSCBC && bc <= kBCInflowTotalNSCBC; }
inline bool bc_is_wall(int bc) { return bc >= kBCWallHeatFlux && bc <= kBCWallInviscid; }

// Input parameters of the path; field meaning follows InputParamClass.
struct CaseParams {
  int n_elem[3] = {0, 0, 0};
  int n_poly_grid = 0, n_poly_sol = 0, n_poly_int_rule = 0;
  int dof_location = 2;        // 1 equidistant, 2 LGL
  int periodic[3] = {1, 1, 1};
  double mach = 0, reynolds = 0, reynolds_length = 1, temp_free_stream = 0;
  double free_stream_dir[3] = {1.0, 0.0, 0.0};   // "Free stream velocity direction"
  double theta_sym = 1, rel_penalty = 1, cfl = 1, dt_synchr = 0;
  int bug_compat = 1;
  // physical faces iMin..kMax (used where periodic[f/2] == 0)
  int bc[6] = {0, 0, 0, 0, 0, 0};
  double bc_data[6][5] = {};   // dimensional prescribed data
  // "NSCBC tuning parameters"
  double outlet_len = 1, outlet_sigma = 0, outlet_beta_l = 1, outlet_beta_t = 1;
  int outlet_one_minus_beta = 1, outlet_average_local = 0;
  double inlet_len = 1, inlet_sigma = 1;
  int inlet_incoming = 1;
  // "Sponge layer", "Characteristic matching layer", sponge block
  int sponge_layer = 0, cml = 0;
  int sponge_elems[6] = {0, 0, 0, 0, 0, 0};
  double damping_constant[6] = {}, damping_exponent[6] = {};
  int monitor_forces = 0;         // "Monitor forces"
  int compute_average = 0;        // "Compute average solution"
  double reference_area = 1.0;    // "Reference area force coefficients"
  int sgs_model = 0;              // "Subgrid scale model": 0 None, 1 WALE, 2 Vreman
  double body_force[3] = {0.0, 0.0, 0.0};   // "Body force per unit volume (x,y,z)", dimensional
  int body_force_set = 0;
};

// Subgrid scale models (InputParamClass.cpp:241-250).
enum : int { kSGSNone = 0, kSGSWale = 1, kSGSVreman = 2 };

struct Tables {
  int P = 0, K = 0, M = 0, Kg = 0;
  std::vector<double> r_int, w_int, r_dof, r_grid;
  // row-major (n_out x n_in) operators
  std::vector<double> A, D;       // M x K: phi_q(r_p), phi'_q(r_p)
  std::vector<double> AT, DT;     // K x M
  std::vector<double> LD;         // K x K: phi_q(r_dof_p)
  std::vector<double> DLD;        // K x K: phi'_q(r_dof_p)
  std::vector<double> VI;         // inverse of LD
  std::vector<double> f_min, f_max, df_min, df_max;  // K
  std::vector<double> w3, w2, basis_center;
  std::vector<double> c0;         // K: phi_q(0)
  // open boundaries / sponge
  std::vector<double> Lq;         // M x K: nodal DOF Lagrange basis at the Gauss points
  std::vector<double> dLq;        // M x M: Gauss-point Lagrange derivative, [point][basis]
  std::vector<double> P1;         // M x M: orthogonal projector onto degree <= P on the Gauss points
  double dell0 = 0;               // -l'_0(r_0) of the nodal DOF basis
};

struct Geometry {
  // per element, as the reference stores it
  std::vector<double> vol;        // [nElem][13][nInt]   J, dr/dx..dt/dz, x,y,z
  std::vector<double> face[6];    // [nElem][16][nInt2D] iMin,iMax,jMin,jMax,kMin,kMax
  std::vector<double> lens;       // [nElem][4] len I, J, K, volume
  std::vector<double> len_les;    // [nElem] LES length scale volume^(1/3) / P
  std::vector<double> nodal;      // [nElem][10][nDOF] J, dr/dx..dt/dz at the nodal
                                  // solution DOFs (SGS model only)
  std::vector<double> ave_norm;   // [nElem][9]
  std::vector<int> nbr;           // [nElem][6], -1 physical boundary
};

// fallback: pRef and rhoRef came from the free-stream data of the parameter file
struct RefConditions { double p_ref, rho_ref, u_ref, mu; int fallback; };

// PC-NSCBC faces of the block.
struct OpenBoundary {
  bool any = false;
  int ig[6][3] = {};              // parametric slots: normal, first, second transverse
  double data[6][5] = {};         // non-dimensional prescribed data
  double area[6] = {};            // boundary surface area
  std::vector<int> faces;         // [nFaces][2]: element, block face; grouped by face
  int begin[7] = {};              // faces of block face f: [begin[f], begin[f+1])
};

// Sponge layer / characteristic matching layer elements.
struct Sponge {
  int n = 0;
  std::vector<int> idx;           // [nElem]: slot or -1
  std::vector<double> fn;         // [n][3][nInt] damping functions
  std::vector<double> coef;       // [n][8]: CML scale per direction, layer sign per direction
};

int n_int1d(int rule);
void build_tables(const CaseParams &p, Tables &t);
// Consistency of the boundary, NSCBC and sponge parameters ("" if valid).
std::string check_case(const CaseParams &p);
// Computes metrics from per-element grid DOFs [nElem][3][Kg^3]; returns an
// error message or "" on success. all = false: the volume and face metric
// terms only of the elements the open-boundary and sponge set-ups read (those
// with a physical face or in a sponge layer; the arrays stay empty when there
// are none), for a caller that computes the rest on the device
// (launch_geometry); the spacing check, the grid nodes, the neighbours and
// the SGS nodal metric as always.
std::string build_geometry(const CaseParams &p, Tables &t, const double *elem_grid, Geometry &g, bool all = true);
// Device table block of launch_geometry (after build_geometry): Lg, dLg
// [M][Kg], lf, dlf [2][Kg], w2 [M^2], w3 [M^3].
std::vector<double> geometry_tables(const Tables &t);
RefConditions reference_conditions(const CaseParams &p);
std::string setup_open_boundary(const CaseParams &p, const Tables &t, const Geometry &g,
                                const RefConditions &rc, OpenBoundary &ob);
std::string setup_sponge(const CaseParams &p, const Tables &t, const Geometry &g, const double *elem_grid,
                         Sponge &sp);
// Averaged nodal primitive state of one element (dimensional, [nDOFs][5])
// -> conservative non-dimensional damping target at the points, [5][nInt].
void damping_state(const Tables &t, const RefConditions &rc, const double *nodal_prim, double *out);

// Nodal dimensional conservative [nDOFs][5] of one element -> modal entropy
// variables [5][nDOFs] (SolverClass::ReadRestartSolution).
void nodal_cons_to_modal_entropy(const Tables &t, const RefConditions &rc,
                                 const double *nodal, double *modal);
// Modal entropy [5][nDOFs] -> nodal dimensional conservative [nDOFs][5].
void modal_entropy_to_nodal_cons(const Tables &t, const RefConditions &rc,
                                 const double *modal, double *nodal);

}  // namespace vcp
