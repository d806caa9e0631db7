// GMRES-polynomial preconditioner roots (host).
#pragma once
#include <complex>
#include <vector>

namespace opem {

// Eigenvalues of a complex upper Hessenberg matrix (row-major n x n, n <= 64) by shifted QR with Givens
// rotations.  Returns false if the iteration does not converge.
bool hessenberg_eigenvalues(int n, std::vector<std::complex<double>> H, std::vector<std::complex<double>> &ev);

// Roots of the degree-d GMRES residual polynomial from d Arnoldi steps: the harmonic Ritz values, the
// eigenvalues of H_d + |h_{d+1,d}|^2 H_d^{-H} e_d e_d^T, in modified Leja order (largest modulus first, then
// each next root maximising the product of distances to the ones chosen).  Hd: row-major d x d upper
// Hessenberg.  Returns false (theta untouched) if H_d is singular or the eigenvalues do not converge.
bool gmres_poly_roots(int d, const std::vector<std::complex<double>> &Hd, double h_next,
                      std::vector<std::complex<double>> &theta);

// Stability of the product form (Loe & Morgan): a root theta_k whose product of other factors
// pof_k = prod_{i != k} |1 - theta_k / theta_i| exceeds 1e4 gets ceil((log10 pof_k - 4) / 14) extra copies,
// appended after the Leja-ordered roots (largest pof first).  Returns the number of copies added.
int poly_add_roots(std::vector<std::complex<double>> &theta);

}  // namespace opem
