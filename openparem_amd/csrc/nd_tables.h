// Nedelec ND_p (p = 1..3) reference tables for affine tetrahedra and triangles.
//
// Basis: phi = lambda^alpha (lambda_i grad lambda_j - lambda_j grad lambda_i), i < j, |alpha| = p-1,
// alpha_m = 0 for m < i, with element vertices in increasing global id (Arnold, Falk, Winther,
// CMAME 2009, Thm 4.4).  This is the space MFEM's ND_FECollection(p) spans
// (src/OpenParEM3D/fem3D.cpp:65-77); the element-local order is 6 edges, 4 faces, cell.
//
// For an affine map x = v0 + J xhat every element integral is a contraction of a handful of
// geometric factors with fixed reference tensors, integrated exactly here:
//   M_e = |det J| sum_kl G_kl  Mhat^{kl},   G = J^-1 J^-T      (6 symmetric components)
//   K_e = 1/|det J| sum_kl C_kl Khat^{kl},  C = J^T J          (6 symmetric components)
//   Q_e = sign(det J) Qhat,   Qhat[a][b] = int phihat_a . curlhat phihat_b
// and for boundary triangles (J2 = 3x2, C2 = J2^T J2):
//   Mf = sqrt(det C2) sum_kl (C2^-1)_kl M2hat^{kl},   Kf = K2hat / sqrt(det C2).
#pragma once
#include <array>
#include <cstdint>
#include <vector>

namespace opem {

struct BasisFn {
  int i, j;          // Whitney pair, i < j (element-local vertex ids)
  int alpha[4];      // lambda exponents
};

struct EntityCounts {
  int edge, face, cell;
};

EntityCounts nd_entity_counts(int p);
int nd_tet_ndof(int p);
int nd_tri_ndof(int p);
std::vector<BasisFn> nd_tet_basis(int p);
std::vector<BasisFn> nd_tri_basis(int p);

// Reference tensors, row-major [comp][a][b].
// tet: comps ordered (00,11,22,01,02,12); off-diagonal comps already hold hat^{kl}+hat^{lk}.
struct TetTables {
  int p = 0, nd = 0;
  std::vector<double> M;   // 6*nd*nd
  std::vector<double> K;   // 6*nd*nd
  std::vector<double> Q;   // nd*nd
  std::vector<double> G;   // nd*4: coefficients of grad(lambda_k) in the basis (affine invariant)
};
struct TriTables {
  int p = 0, nd = 0;
  std::vector<double> M;   // 3*nd*nd  comps (00,11,01)
  std::vector<double> K;   // nd*nd
};

const TetTables &tet_tables(int p);
const TriTables &tri_tables(int p);

// Geometric factors of one tet (vertices in element-local = sorted order).
// gm[6] multiplies M comps, gk[6] multiplies K comps; returns det J.
double tet_factors(const double v[4][3], double gm[6], double gk[6]);
// Boundary triangle: fm[3] multiplies M2 comps; fk multiplies K2.
void tri_factors(const double v[3][3], double fm[3], double *fk);

}  // namespace opem
