// Roots of the GMRES polynomial preconditioner (Loe & Morgan, "Toward efficient polynomial preconditioning
// for GMRES", Numer. Linear Algebra Appl. 29 (2022)): harmonic Ritz values of a short Arnoldi run, Leja order.
#include "poly.h"

#include <algorithm>
#include <cmath>

namespace opem {

using cplx = std::complex<double>;

bool hessenberg_eigenvalues(int n, std::vector<cplx> H, std::vector<cplx> &ev) {
  ev.assign(n, cplx(0.0));
  auto at = [&](int i, int j) -> cplx & { return H[(size_t)i * n + j]; };
  int hi = n - 1, iter = 0;
  std::vector<cplx> cs(n), sn(n);
  while (hi >= 0) {
    int lo = hi;
    while (lo > 0) {
      const double sub = std::abs(at(lo, lo - 1));
      double dsum = std::abs(at(lo, lo)) + std::abs(at(lo - 1, lo - 1));
      if (dsum == 0.0) dsum = 1.0;
      if (sub <= 1e-15 * dsum) {
        at(lo, lo - 1) = 0.0;
        break;
      }
      lo--;
    }
    if (lo == hi) {
      ev[hi] = at(hi, hi);
      hi--;
      iter = 0;
      continue;
    }
    if (++iter > 60 * n) return false;
    // Wilkinson shift: the eigenvalue of the trailing 2 x 2 block closer to its last diagonal entry
    const cplx a = at(hi - 1, hi - 1), b = at(hi - 1, hi), c = at(hi, hi - 1), d = at(hi, hi);
    const cplx tr = a + d, det = a * d - b * c;
    const cplx disc = std::sqrt(tr * tr / 4.0 - det);
    const cplx l1 = tr / 2.0 + disc, l2 = tr / 2.0 - disc;
    cplx mu = std::abs(l1 - d) < std::abs(l2 - d) ? l1 : l2;
    if (iter % 11 == 10) mu += std::abs(c);  // exceptional shift
    for (int k = lo; k <= hi; k++) at(k, k) -= mu;
    for (int k = lo; k < hi; k++) {  // H - mu I = Q R (rows), then R Q
      const cplx x = at(k, k), y = at(k + 1, k);
      const double r = std::sqrt(std::norm(x) + std::norm(y));
      cs[k] = r > 0 ? x / r : cplx(1.0);
      sn[k] = r > 0 ? y / r : cplx(0.0);
      for (int j = k; j <= hi; j++) {
        const cplx t1 = at(k, j), t2 = at(k + 1, j);
        at(k, j) = std::conj(cs[k]) * t1 + std::conj(sn[k]) * t2;
        at(k + 1, j) = -sn[k] * t1 + cs[k] * t2;
      }
    }
    for (int k = lo; k < hi; k++) {
      for (int i = lo; i <= std::min(k + 1, hi); i++) {
        const cplx t1 = at(i, k), t2 = at(i, k + 1);
        at(i, k) = t1 * cs[k] + t2 * sn[k];
        at(i, k + 1) = -t1 * std::conj(sn[k]) + t2 * std::conj(cs[k]);
      }
    }
    for (int k = lo; k <= hi; k++) at(k, k) += mu;
  }
  return true;
}

bool gmres_poly_roots(int d, const std::vector<cplx> &Hd, double h_next, std::vector<cplx> &theta) {
  if (d < 1) return false;
  // f = H_d^{-H} e_d: solve H_d^H f = e_d by Gaussian elimination with partial pivoting
  std::vector<cplx> A((size_t)d * d), f(d, cplx(0.0));
  for (int i = 0; i < d; i++)
    for (int j = 0; j < d; j++) A[(size_t)i * d + j] = std::conj(Hd[(size_t)j * d + i]);
  f[d - 1] = 1.0;
  for (int k = 0; k < d; k++) {
    int p = k;
    for (int i = k + 1; i < d; i++)
      if (std::abs(A[(size_t)i * d + k]) > std::abs(A[(size_t)p * d + k])) p = i;
    if (std::abs(A[(size_t)p * d + k]) == 0.0) return false;
    if (p != k) {
      for (int j = 0; j < d; j++) std::swap(A[(size_t)k * d + j], A[(size_t)p * d + j]);
      std::swap(f[k], f[p]);
    }
    for (int i = k + 1; i < d; i++) {
      const cplx m = A[(size_t)i * d + k] / A[(size_t)k * d + k];
      for (int j = k; j < d; j++) A[(size_t)i * d + j] -= m * A[(size_t)k * d + j];
      f[i] -= m * f[k];
    }
  }
  for (int i = d - 1; i >= 0; i--) {
    cplx t = f[i];
    for (int j = i + 1; j < d; j++) t -= A[(size_t)i * d + j] * f[j];
    f[i] = t / A[(size_t)i * d + i];
  }
  std::vector<cplx> G(Hd);
  for (int i = 0; i < d; i++) G[(size_t)i * d + d - 1] += h_next * h_next * f[i];
  std::vector<cplx> ev;
  if (!hessenberg_eigenvalues(d, G, ev)) return false;
  for (const cplx &e : ev)
    if (!std::isfinite(e.real()) || !std::isfinite(e.imag()) || std::abs(e) == 0.0) return false;
  // modified Leja order
  std::vector<int> used(d, 0);
  theta.clear();
  int first = 0;
  for (int i = 1; i < d; i++)
    if (std::abs(ev[i]) > std::abs(ev[first])) first = i;
  used[first] = 1;
  theta.push_back(ev[first]);
  std::vector<double> logprod(d, 0.0);
  for (int step = 1; step < d; step++) {
    int best = -1;
    double bv = -1e300;
    for (int i = 0; i < d; i++) {
      if (used[i]) continue;
      logprod[i] += std::log(std::abs(ev[i] - theta.back()) + 1e-300);
      if (logprod[i] > bv) {
        bv = logprod[i];
        best = i;
      }
    }
    used[best] = 1;
    theta.push_back(ev[best]);
  }
  return true;
}

int poly_add_roots(std::vector<cplx> &theta) {
  const int d = (int)theta.size();
  std::vector<std::pair<double, int>> big;
  for (int k = 0; k < d; k++) {
    double lp = 0.0;
    for (int i = 0; i < d; i++)
      if (i != k) lp += std::log10(std::abs(1.0 - theta[k] / theta[i]) + 1e-300);
    if (lp > 4.0) big.push_back({lp, k});
  }
  std::sort(big.begin(), big.end(), [](const std::pair<double, int> &a, const std::pair<double, int> &b) {
    return a.first > b.first;
  });
  int added = 0;
  for (const auto &e : big) {
    const int copies = (int)std::ceil((e.first - 4.0) / 14.0);
    for (int c = 0; c < copies; c++) theta.push_back(theta[e.second]);
    added += copies;
  }
  return added;
}

}  // namespace opem
