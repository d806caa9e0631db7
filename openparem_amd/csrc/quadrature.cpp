// Golub-Welsch Gauss-Jacobi rules (symmetric tridiagonal Jacobi matrix, cyclic Jacobi eigensolver).
#include "quadrature.h"

#include <cmath>
#include <stdexcept>

namespace opem {

void gauss_jacobi01(int n, double alpha, std::vector<double> &t, std::vector<double> &w) {
  if (n < 1 || n > 32) throw std::runtime_error("gauss_jacobi01: n out of range");
  const double a = alpha, b = 0.0;
  std::vector<double> J(n * n, 0.0);
  for (int k = 0; k < n; k++) {
    const double s = 2.0 * k + a + b;
    double d;
    if (k == 0)
      d = (b - a) / (a + b + 2.0);
    else
      d = (b * b - a * a) / (s * (s + 2.0));
    J[k * n + k] = d;
    if (k >= 1) {
      const double num = 4.0 * k * (k + a) * (k + b) * (k + a + b);
      const double den = s * s * (s + 1.0) * (s - 1.0);
      const double off = std::sqrt(num / den);
      J[k * n + k - 1] = J[(k - 1) * n + k] = off;
    }
  }
  // eigenvectors by cyclic Jacobi rotations
  std::vector<double> V(n * n, 0.0);
  for (int i = 0; i < n; i++) V[i * n + i] = 1.0;
  for (int sweep = 0; sweep < 100; sweep++) {
    double off = 0.0;
    for (int p = 0; p < n; p++)
      for (int q = p + 1; q < n; q++) off += J[p * n + q] * J[p * n + q];
    if (off < 1e-34) break;
    for (int p = 0; p < n; p++)
      for (int q = p + 1; q < n; q++) {
        const double apq = J[p * n + q];
        if (std::fabs(apq) < 1e-300) continue;
        const double theta = (J[q * n + q] - J[p * n + p]) / (2.0 * apq);
        const double tt = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(tt * tt + 1.0), s = tt * c;
        for (int k = 0; k < n; k++) {
          const double jkp = J[k * n + p], jkq = J[k * n + q];
          J[k * n + p] = c * jkp - s * jkq;
          J[k * n + q] = s * jkp + c * jkq;
        }
        for (int k = 0; k < n; k++) {
          const double jpk = J[p * n + k], jqk = J[q * n + k];
          J[p * n + k] = c * jpk - s * jqk;
          J[q * n + k] = s * jpk + c * jqk;
        }
        for (int k = 0; k < n; k++) {
          const double vkp = V[k * n + p], vkq = V[k * n + q];
          V[k * n + p] = c * vkp - s * vkq;
          V[k * n + q] = s * vkp + c * vkq;
        }
      }
  }
  // mu0 = int_{-1}^{1} (1-x)^a dx = 2^{a+1}/(a+1)
  const double mu0 = std::pow(2.0, a + 1.0) / (a + 1.0);
  std::vector<int> idx(n);
  for (int i = 0; i < n; i++) idx[i] = i;
  for (int i = 0; i < n; i++)
    for (int j = i + 1; j < n; j++)
      if (J[idx[j] * n + idx[j]] < J[idx[i] * n + idx[i]]) std::swap(idx[i], idx[j]);
  t.resize(n);
  w.resize(n);
  for (int i = 0; i < n; i++) {
    const int k = idx[i];
    const double x = J[k * n + k];
    const double v0 = V[0 * n + k];
    t[i] = 0.5 * (x + 1.0);
    w[i] = mu0 * v0 * v0 / std::pow(2.0, a + 1.0);
  }
}

void tri_rule(int n, std::vector<double> &L, std::vector<double> &w) {
  std::vector<double> u, wu, s, ws;
  gauss_jacobi01(n, 1.0, u, wu);
  gauss_jacobi01(n, 0.0, s, ws);
  L.clear();
  w.clear();
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) {
      const double x = u[i], y = s[j] * (1.0 - u[i]);
      L.push_back(1.0 - x - y);
      L.push_back(x);
      L.push_back(y);
      w.push_back(wu[i] * ws[j]);
    }
}

void tet_rule(int n, std::vector<double> &L, std::vector<double> &w) {
  std::vector<double> u, wu, v, wv, s, ws;
  gauss_jacobi01(n, 2.0, u, wu);
  gauss_jacobi01(n, 1.0, v, wv);
  gauss_jacobi01(n, 0.0, s, ws);
  L.clear();
  w.clear();
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++)
      for (int k = 0; k < n; k++) {
        const double x = u[i], y = v[j] * (1.0 - u[i]), z = s[k] * (1.0 - u[i]) * (1.0 - v[j]);
        L.push_back(1.0 - x - y - z);
        L.push_back(x);
        L.push_back(y);
        L.push_back(z);
        w.push_back(wu[i] * wv[j] * ws[k]);
      }
}

}  // namespace opem
