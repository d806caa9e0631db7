// Collapsed (Duffy) Gauss-Jacobi rules on the unit triangle and tetrahedron, nodes from Golub-Welsch.
#pragma once
#include <vector>

namespace opem {

// Gauss-Jacobi on [0,1] with weight (1-t)^alpha (beta = 0): n nodes and weights.
void gauss_jacobi01(int n, double alpha, std::vector<double> &t, std::vector<double> &w);

// Triangle rule with n*n points: barycentrics L (3 per point) and weights (sum 1/2).
// u ~ GJ(alpha=1), s ~ GJ(alpha=0), x = u, y = s(1-u), L = (1-x-y, x, y); u outer loop.
void tri_rule(int n, std::vector<double> &L, std::vector<double> &w);

// Tetrahedron rule with n^3 points: barycentrics L (4 per point) and weights (sum 1/6).
// u ~ GJ(alpha=2), v ~ GJ(alpha=1), s ~ GJ(alpha=0); x = u, y = v(1-u), z = s(1-u)(1-v); u outer, s inner.
void tet_rule(int n, std::vector<double> &L, std::vector<double> &w);

}  // namespace opem
