// Smoothed-aggregation gradient coarse space for the two-level Schwarz preconditioner.
//
// C = G P : G embeds vertex hat-function gradients into ND_p (grad lambda_v expressed in the element
// basis, an affine-invariant table), P aggregates free vertices into spatial bins and smooths the
// aggregate indicators with one damped-Jacobi step of the P1 Laplacian.  Vertices on Dirichlet-type
// boundaries (PEC and every port) are excluded, so C has no rows on any excitation's Dirichlet DOFs
// and C^T A C is the same for every excitation of a frequency.  Because A(omega) = sum_k c_k A_k,
// the coarse matrix is sum_k c_k (C^T A_k C): the C^T A_k C are formed once per mesh.
#pragma once
#include <cstdint>
#include <vector>

#include "space.h"

namespace opem {

struct Coarse {
  int nc = 0;
  std::vector<int32_t> row_ptr, row_col;  // C by DOF rows
  std::vector<double> row_val;
  std::vector<int32_t> col_ptr, col_row;  // C^T by coarse rows
  std::vector<double> col_val;
};

// vertex_excluded[v] != 0 removes vertex v; target_nc bounds the number of aggregates (<= 128).
void build_coarse(const Mesh &m, const Space &s, const std::vector<uint8_t> &vertex_excluded, int target_nc,
                  Coarse &c);

// Ac[a][b] = sum_{(i,j) in pattern} C_ia vals[nz(i,j)] C_jb  for one real matrix on the pattern
// (vals == nullptr with (ov_nz, ov_val, nov) given: the matrix is a sparse overlay on the pattern).
void coarse_galerkin(const Space &s, const Coarse &c, const double *vals, const int32_t *ov_nz,
                     const double *ov_val, int64_t nov, std::vector<double> &Ac, int nthreads);

}  // namespace opem
