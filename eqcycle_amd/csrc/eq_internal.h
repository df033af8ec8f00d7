// Internal declarations shared by the eqcycle HIP library sources.
//
// Device layout: nodal fields live on the structured GLL grid, node (I, J) at
// n = J * NX + I, with I = 0..4*NelX along dip (x) and J = 0..4*NelY fault
// normal (y).  The fault is row J = 0 and fault node j (reference order, deep
// to shallow) is I = j.  The reference's element-by-element numbering
// (src/MeshBox.jl) is converted once, at the C-ABI boundary.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/eqcycle_hip.h"

namespace eq {

// Geometry and operator coefficients passed by value to every grid kernel.
struct Geom {
  int NX, NY;         // nodes per direction
  int nelx, nely;
  double c1, c2;      // jac/dx_dxi^2 (x term), jac/dy_deta^2 (y term)
  double D[25];       // D[p*5+i] = h'_i(xi_p)  (= H[i,p])
  double w[5];        // GLL weights
  double x[5];        // GLL nodes
};

// Host-side RS-AMG hierarchy (CSR, int32 columns), built by amg_build().
struct HostCsr {
  int64_t n = 0, m = 0;
  std::vector<int64_t> rowptr;
  std::vector<int32_t> col;
  std::vector<double> val;
  int64_t nnz() const { return rowptr.empty() ? 0 : rowptr[n]; }
};

struct HostLevel {
  HostCsr A, P, R;    // R = P^T
  std::vector<double> dinv;
  double lambda_max = 0;  // of D^-1 A
};

struct AmgOptions {
  double theta = 0.25;
  int max_levels = 10;
  int64_t max_coarse = 10;
  int strength_abs = 1;   // 1: |a_ij| >= theta max|a_ik| (AlgebraicMultigrid.jl Classical)
  int build_on_negated = 1;  // build from kni = -K_II as the reference does
};

struct HostHierarchy {
  std::vector<HostLevel> lev;   // lev[0].A = fine operator
  std::vector<double> coarse_inv;  // dense inverse of the coarsest A (row-major)
  int64_t ncoarse = 0;
};

// amg_host.cpp
// splits (optional): the C/F split of each level that has a coarser one, 1 = C
int amg_build(const HostCsr& A, const AmgOptions& opt, HostHierarchy* out, std::string* err,
              std::vector<std::vector<int>>* splits = nullptr);
// D^-1, lambda_max(D^-1 A) (omega: the caller's Jacobi weights per level, or
// null to estimate) and the dense coarsest inverse, once every level's A, P
// and R are set.
int finish_hierarchy(HostHierarchy* H, const double* omega, std::string* err);
void csr_transpose(const HostCsr& A, HostCsr* T);
double estimate_lambda_max_dinvA(const HostCsr& A, const std::vector<double>& dinv, int iters);
// K_II in structured interior order (reduced index r = n - NX), from nodal mu
void assemble_kii(const Geom& g, const double* mu_struct, HostCsr* K);
// Condensed sweeps of a coarse level with one Jacobi step (weight omega, zero
// start) before and after the coarse-grid correction:
//   b_next = M b,            M = R - omega R A D^-1
//   e      = U b + V e_next, U = 2 omega D^-1 - omega^2 D^-1 A D^-1,
//                            V = P - omega D^-1 A P
// the same linear operators as smoothing, residual, restriction,
// prolongation and smoothing applied one after the other.
void condense_level(const HostLevel& L, double omega, HostCsr* M, HostCsr* U, HostCsr* V);
// The condensed levels 1 .. k (M[j - 1], U[j - 1], V[j - 1]) above a dense level
// k + 1, collapsed to one product down and one up:
//   [b_2; ...; b_(k+1)] = S b_1,                 S = [M_1; M_2 M_1; ...; M_k ... M_1]
//   e_1 = C [b_1, ..., b_k, e_(k+1)],            C = [U_1, V_1 U_2, ..., V_1 ... V_(k-1) U_k, V_1 ... V_k]
// (e_(k+1) the dense level's correction); C's column block j starts at
// col_off[j] of a vector of ncols entries (block k: e_(k+1)).
void collapse_coarse(const std::vector<HostCsr>& M, const std::vector<HostCsr>& U, const std::vector<HostCsr>& V,
                     const std::vector<int64_t>& col_off, int64_t ncols, HostCsr* S, HostCsr* C);
// p-multigrid: level 0 = K_II with the Q4 -> Q1 vertex prolongation, then RS on
// the Galerkin Q1 operator.
void build_pmg_p0(const Geom& g, HostCsr* P);
int pmg_build(const Geom& g, const HostCsr& Kii, const AmgOptions& opt, HostHierarchy* H, std::string* err);

}  // namespace eq
