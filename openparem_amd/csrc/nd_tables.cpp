// Exact reference tensors for the ND_p basis (see nd_tables.h).
#include "nd_tables.h"

#include <cmath>
#include <map>
#include <mutex>
#include <stdexcept>

namespace opem {
namespace {

// Non-decreasing index tuples of length t over [0,n), lexicographic (itertools order).
void multi_indices(int n, int t, std::vector<std::array<int, 4>> &out) {
  std::vector<int> c(t, 0);
  auto emit = [&]() {
    std::array<int, 4> a{0, 0, 0, 0};
    for (int k : c) a[k]++;
    out.push_back(a);
  };
  if (t == 0) {
    emit();
    return;
  }
  while (true) {
    emit();
    int k = t - 1;
    while (k >= 0 && c[k] == n - 1) k--;
    if (k < 0) break;
    c[k]++;
    for (int m = k + 1; m < t; m++) c[m] = c[k];
  }
}

// Patterns owned by a sub-simplex with nv vertices, in its own vertex positions.
std::vector<BasisFn> entity_patterns(int p, int nv) {
  std::vector<BasisFn> pats;
  std::vector<std::array<int, 4>> alphas;
  multi_indices(nv, p - 1, alphas);
  for (int i = 0; i < nv; i++)
    for (int j = i + 1; j < nv; j++)
      for (auto &a : alphas) {
        bool ok = true;
        for (int m = 0; m < i; m++)
          if (a[m]) ok = false;
        if (!ok) continue;
        int supp = 0;
        for (int m = 0; m < nv; m++)
          if (a[m] || m == i || m == j) supp++;
        if (supp != nv) continue;
        BasisFn f{i, j, {a[0], a[1], a[2], a[3]}};
        pats.push_back(f);
      }
  return pats;
}

void lift(const int *ent, int ne, const std::vector<BasisFn> &pats, std::vector<BasisFn> &out) {
  for (auto &pt : pats) {
    BasisFn f{ent[pt.i], ent[pt.j], {0, 0, 0, 0}};
    for (int k = 0; k < ne; k++) f.alpha[ent[k]] += pt.alpha[k];
    out.push_back(f);
  }
}

const int TET_E[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
const int TET_F[4][3] = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}};
const int TRI_E[3][2] = {{0, 1}, {0, 2}, {1, 2}};

// Vector polynomial: sum of coef[3] * lambda^beta.
struct Term {
  double c[3];
  int b[4];
};
using VPoly = std::vector<Term>;

double fact(int n) {
  double r = 1;
  for (int k = 2; k <= n; k++) r *= k;
  return r;
}
// int_{ref simplex} lambda^beta ; nvars = 4 (tet, vol 1/6) or 3 (tri, area 1/2)
double mono_int(const int *b, int nvars) {
  double num = 1;
  int s = 0;
  for (int k = 0; k < nvars; k++) {
    num *= fact(b[k]);
    s += b[k];
  }
  return num / fact(s + nvars - 1);
}

void cross(const double *a, const double *b, double *c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

// phi and curl of one basis function as vector polynomials, given reference gradients g[nv][3].
void basis_polys(const BasisFn &f, const double g[4][3], VPoly &phi, VPoly &curl) {
  phi.clear();
  curl.clear();
  Term t{};
  for (int k = 0; k < 4; k++) t.b[k] = f.alpha[k];
  t.b[f.i]++;
  for (int d = 0; d < 3; d++) t.c[d] = g[f.j][d];
  phi.push_back(t);
  Term u{};
  for (int k = 0; k < 4; k++) u.b[k] = f.alpha[k];
  u.b[f.j]++;
  for (int d = 0; d < 3; d++) u.c[d] = -g[f.i][d];
  phi.push_back(u);
  // curl = sum_m alpha_m lambda^{alpha-e_m} g_m x (lambda_i g_j - lambda_j g_i) + 2 lambda^alpha g_i x g_j
  for (int m = 0; m < 4; m++) {
    if (!f.alpha[m]) continue;
    double gmj[3], gmi[3];
    cross(g[m], g[f.j], gmj);
    cross(g[m], g[f.i], gmi);
    Term a{}, b{};
    for (int k = 0; k < 4; k++) a.b[k] = b.b[k] = f.alpha[k];
    a.b[m]--;
    b.b[m]--;
    a.b[f.i]++;
    b.b[f.j]++;
    for (int d = 0; d < 3; d++) {
      a.c[d] = f.alpha[m] * gmj[d];
      b.c[d] = -f.alpha[m] * gmi[d];
    }
    curl.push_back(a);
    curl.push_back(b);
  }
  double gij[3];
  cross(g[f.i], g[f.j], gij);
  Term w{};
  for (int k = 0; k < 4; k++) w.b[k] = f.alpha[k];
  for (int d = 0; d < 3; d++) w.c[d] = 2.0 * gij[d];
  curl.push_back(w);
}

// int P_k * Q_l over the reference simplex.
double prod_int(const VPoly &P, int k, const VPoly &Q, int l, int nvars) {
  double s = 0;
  for (auto &a : P)
    for (auto &b : Q) {
      double c = a.c[k] * b.c[l];
      if (c == 0) continue;
      int bb[4];
      for (int m = 0; m < 4; m++) bb[m] = a.b[m] + b.b[m];
      s += c * mono_int(bb, nvars);
    }
  return s;
}

TetTables build_tet(int p) {
  TetTables T;
  T.p = p;
  auto basis = nd_tet_basis(p);
  int nd = (int)basis.size();
  T.nd = nd;
  const double g[4][3] = {{-1, -1, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  std::vector<VPoly> phi(nd), curl(nd);
  for (int a = 0; a < nd; a++) basis_polys(basis[a], g, phi[a], curl[a]);
  const int comp[6][2] = {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {0, 2}, {1, 2}};
  T.M.assign(6 * nd * nd, 0.0);
  T.K.assign(6 * nd * nd, 0.0);
  T.Q.assign(nd * nd, 0.0);
  for (int c = 0; c < 6; c++) {
    int k = comp[c][0], l = comp[c][1];
    for (int a = 0; a < nd; a++)
      for (int b = 0; b < nd; b++) {
        double m = prod_int(phi[a], k, phi[b], l, 4);
        double q = prod_int(curl[a], k, curl[b], l, 4);
        if (k != l) {
          m += prod_int(phi[a], l, phi[b], k, 4);
          q += prod_int(curl[a], l, curl[b], k, 4);
        }
        T.M[(c * nd + a) * nd + b] = m;
        T.K[(c * nd + a) * nd + b] = q;
      }
  }
  for (int a = 0; a < nd; a++)
    for (int b = 0; b < nd; b++) {
      double s = 0;
      for (int k = 0; k < 3; k++) s += prod_int(phi[a], k, curl[b], k, 4);
      T.Q[a * nd + b] = s;
    }
  // grad(lambda_k) = sum_a G[a][k] phi_a: solve Mref G = int phi . ghat_k on the reference tet
  std::vector<double> Mr(nd * nd), rhs(nd * 4, 0.0);
  for (int a = 0; a < nd; a++)
    for (int b = 0; b < nd; b++) Mr[a * nd + b] = T.M[(0 * nd + a) * nd + b] + T.M[(1 * nd + a) * nd + b] +
                                                  T.M[(2 * nd + a) * nd + b];
  for (int a = 0; a < nd; a++)
    for (auto &t : phi[a]) {
      const double w = mono_int(t.b, 4);
      for (int k = 0; k < 4; k++) rhs[a * 4 + k] += w * (t.c[0] * g[k][0] + t.c[1] * g[k][1] + t.c[2] * g[k][2]);
    }
  // Gaussian elimination with partial pivoting (nd <= 45, 4 right-hand sides)
  for (int k = 0; k < nd; k++) {
    int p = k;
    for (int i = k + 1; i < nd; i++)
      if (std::fabs(Mr[i * nd + k]) > std::fabs(Mr[p * nd + k])) p = i;
    if (p != k) {
      for (int c = 0; c < nd; c++) std::swap(Mr[k * nd + c], Mr[p * nd + c]);
      for (int c = 0; c < 4; c++) std::swap(rhs[k * 4 + c], rhs[p * 4 + c]);
    }
    for (int i = k + 1; i < nd; i++) {
      const double f = Mr[i * nd + k] / Mr[k * nd + k];
      for (int c = k; c < nd; c++) Mr[i * nd + c] -= f * Mr[k * nd + c];
      for (int c = 0; c < 4; c++) rhs[i * 4 + c] -= f * rhs[k * 4 + c];
    }
  }
  T.G.assign(nd * 4, 0.0);
  for (int i = nd - 1; i >= 0; i--)
    for (int c = 0; c < 4; c++) {
      double t = rhs[i * 4 + c];
      for (int j = i + 1; j < nd; j++) t -= Mr[i * nd + j] * T.G[j * 4 + c];
      T.G[i * 4 + c] = t / Mr[i * nd + i];
    }
  for (double &v : T.G)  // entries are small rationals; remove solve round-off
    v = std::round(v * 1e9) / 1e9;
  return T;
}

TriTables build_tri(int p) {
  TriTables T;
  T.p = p;
  auto basis = nd_tri_basis(p);
  int nd = (int)basis.size();
  T.nd = nd;
  const double g[4][3] = {{-1, -1, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 0}};
  std::vector<VPoly> phi(nd), curl(nd);
  for (int a = 0; a < nd; a++) basis_polys(basis[a], g, phi[a], curl[a]);
  const int comp[3][2] = {{0, 0}, {1, 1}, {0, 1}};
  T.M.assign(3 * nd * nd, 0.0);
  T.K.assign(nd * nd, 0.0);
  for (int c = 0; c < 3; c++) {
    int k = comp[c][0], l = comp[c][1];
    for (int a = 0; a < nd; a++)
      for (int b = 0; b < nd; b++) {
        double m = prod_int(phi[a], k, phi[b], l, 3);
        if (k != l) m += prod_int(phi[a], l, phi[b], k, 3);
        T.M[(c * nd + a) * nd + b] = m;
      }
  }
  for (int a = 0; a < nd; a++)
    for (int b = 0; b < nd; b++) T.K[a * nd + b] = prod_int(curl[a], 2, curl[b], 2, 3);
  return T;
}

}  // namespace

EntityCounts nd_entity_counts(int p) {
  return {(int)entity_patterns(p, 2).size(), (int)entity_patterns(p, 3).size(),
          (int)entity_patterns(p, 4).size()};
}

int nd_tet_ndof(int p) { return p * (p + 2) * (p + 3) / 2; }
int nd_tri_ndof(int p) { return p * (p + 2); }

std::vector<BasisFn> nd_tet_basis(int p) {
  if (p < 1 || p > 3) throw std::runtime_error("ND order must be 1..3");
  std::vector<BasisFn> out;
  auto pe = entity_patterns(p, 2), pf = entity_patterns(p, 3), pc = entity_patterns(p, 4);
  for (auto &e : TET_E) lift(e, 2, pe, out);
  for (auto &f : TET_F) lift(f, 3, pf, out);
  const int cell[4] = {0, 1, 2, 3};
  lift(cell, 4, pc, out);
  return out;
}

std::vector<BasisFn> nd_tri_basis(int p) {
  if (p < 1 || p > 3) throw std::runtime_error("ND order must be 1..3");
  std::vector<BasisFn> out;
  auto pe = entity_patterns(p, 2), pf = entity_patterns(p, 3);
  for (auto &e : TRI_E) lift(e, 2, pe, out);
  const int face[3] = {0, 1, 2};
  lift(face, 3, pf, out);
  return out;
}

const TetTables &tet_tables(int p) {
  static std::mutex mu;
  static std::map<int, TetTables> cache;
  std::lock_guard<std::mutex> lk(mu);
  auto it = cache.find(p);
  if (it == cache.end()) it = cache.emplace(p, build_tet(p)).first;
  return it->second;
}

const TriTables &tri_tables(int p) {
  static std::mutex mu;
  static std::map<int, TriTables> cache;
  std::lock_guard<std::mutex> lk(mu);
  auto it = cache.find(p);
  if (it == cache.end()) it = cache.emplace(p, build_tri(p)).first;
  return it->second;
}

double tet_factors(const double v[4][3], double gm[6], double gk[6]) {
  double J[3][3];  // columns v1-v0, v2-v0, v3-v0
  for (int c = 0; c < 3; c++)
    for (int r = 0; r < 3; r++) J[r][c] = v[c + 1][r] - v[0][r];
  double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
               J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
               J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
  double inv[3][3];
  inv[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / det;
  inv[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
  inv[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
  inv[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) / det;
  inv[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
  inv[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
  inv[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / det;
  inv[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
  inv[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
  const int comp[6][2] = {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {0, 2}, {1, 2}};
  double ad = std::fabs(det);
  for (int c = 0; c < 6; c++) {
    int k = comp[c][0], l = comp[c][1];
    double G = 0, C = 0;
    for (int m = 0; m < 3; m++) {
      G += inv[k][m] * inv[l][m];  // (J^-1 J^-T)_kl
      C += J[m][k] * J[m][l];      // (J^T J)_kl
    }
    gm[c] = ad * G;
    gk[c] = C / ad;
  }
  return det;
}

void tri_factors(const double v[3][3], double fm[3], double *fk) {
  double e1[3], e2[3];
  for (int d = 0; d < 3; d++) {
    e1[d] = v[1][d] - v[0][d];
    e2[d] = v[2][d] - v[0][d];
  }
  double c00 = e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2];
  double c11 = e2[0] * e2[0] + e2[1] * e2[1] + e2[2] * e2[2];
  double c01 = e1[0] * e2[0] + e1[1] * e2[1] + e1[2] * e2[2];
  double det = c00 * c11 - c01 * c01;
  double sq = std::sqrt(det);
  fm[0] = sq * (c11 / det);
  fm[1] = sq * (c00 / det);
  fm[2] = sq * (-c01 / det);
  *fk = 1.0 / sq;
}

}  // namespace opem
