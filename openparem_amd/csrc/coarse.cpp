// Smoothed-aggregation gradient coarse space (see coarse.h).
#include "coarse.h"

#include <algorithm>
#include <cmath>
#include <map>
#include <stdexcept>
#include <thread>

#include "nd_tables.h"

namespace opem {

void build_coarse(const Mesh &m, const Space &s, const std::vector<uint8_t> &excl, int target_nc, Coarse &c) {
  c = Coarse();
  if (target_nc < 1 || target_nc > 128) throw std::runtime_error("coarse size must be 1..128");
  const TetTables &T = tet_tables(s.p);
  const int nd = s.nd;
  const int64_t nt = m.nt, nv = m.nv, N = s.ndof;
  std::vector<int32_t> fid(nv, -1);
  int32_t nf = 0;
  for (int64_t v = 0; v < nv; v++)
    if (!excl[v]) fid[v] = nf++;
  if (nf == 0) return;
  // spatial bins over the free vertices, at most target_nc of them
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int64_t v = 0; v < nv; v++)
    if (fid[v] >= 0)
      for (int d = 0; d < 3; d++) {
        lo[d] = std::min(lo[d], m.xyz[v * 3 + d]);
        hi[d] = std::max(hi[d], m.xyz[v * 3 + d]);
      }
  double ext[3], vol = 1.0;
  for (int d = 0; d < 3; d++) {
    ext[d] = std::max(hi[d] - lo[d], 1e-30);
    vol *= ext[d];
  }
  double side = std::cbrt(vol / target_nc);
  int nb[3];
  std::vector<int32_t> agg(nf);
  int nagg = 0;
  for (int tries = 0; tries < 200; tries++) {
    for (int d = 0; d < 3; d++) nb[d] = std::max(1, (int)std::lround(ext[d] / side));
    std::map<int64_t, int> ids;
    for (int64_t v = 0; v < nv; v++) {
      if (fid[v] < 0) continue;
      int64_t key = 0;
      for (int d = 2; d >= 0; d--) {
        int b = (int)std::floor((m.xyz[v * 3 + d] - lo[d]) / ext[d] * nb[d]);
        b = std::min(std::max(b, 0), nb[d] - 1);
        key = key * nb[d] + b;
      }
      auto it = ids.find(key);
      if (it == ids.end()) it = ids.emplace(key, (int)ids.size()).first;
      agg[fid[v]] = it->second;
    }
    nagg = (int)ids.size();
    if (nagg <= target_nc) break;
    side *= 1.05;
  }
  c.nc = nagg;
  // P1 Laplacian rows of free vertices
  std::vector<std::map<int32_t, double>> L(nf);
  for (int64_t e = 0; e < nt; e++) {
    const int32_t *tv = &s.tets_sorted[e * 4];
    double J[3][3];
    for (int a = 0; a < 3; a++)
      for (int d = 0; d < 3; d++) J[d][a] = m.xyz[(int64_t)tv[a + 1] * 3 + d] - m.xyz[(int64_t)tv[0] * 3 + d];
    const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                       J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                       J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    double g[4][3];
    // rows of J^-1 are grad lambda_1..3
    g[1][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / det;
    g[1][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
    g[1][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
    g[2][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) / det;
    g[2][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
    g[2][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
    g[3][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / det;
    g[3][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
    g[3][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
    for (int d = 0; d < 3; d++) g[0][d] = -(g[1][d] + g[2][d] + g[3][d]);
    const double vol6 = std::fabs(det) / 6.0;
    for (int a = 0; a < 4; a++) {
      const int32_t fa = fid[tv[a]];
      if (fa < 0) continue;
      for (int b = 0; b < 4; b++) {
        const int32_t fb = fid[tv[b]];
        if (fb < 0) continue;
        L[fa][fb] += vol6 * (g[a][0] * g[b][0] + g[a][1] * g[b][1] + g[a][2] * g[b][2]);
      }
    }
  }
  // smoothed prolongator rows: P = (I - w D^-1 L) P_agg
  const double w = 2.0 / 3.0;
  std::vector<std::map<int32_t, double>> P(nf);
  for (int32_t v = 0; v < nf; v++) {
    const double dv = L[v].count(v) ? L[v][v] : 1.0;
    P[v][agg[v]] += 1.0;
    for (auto &kv : L[v]) P[v][agg[kv.first]] -= w * kv.second / dv;
  }
  // C rows: sum over the DOF's gradient coefficients (owner element; conformity makes them unique)
  c.row_ptr.assign(N + 1, 0);
  std::vector<std::map<int32_t, double>> crow(N);
  for (int64_t d = 0; d < N; d++) {
    const int32_t e = s.owner_elem[d], a = s.owner_local[d];
    for (int k = 0; k < 4; k++) {
      const double gv = T.G[a * 4 + k];
      if (gv == 0.0) continue;
      const int32_t fv = fid[s.tets_sorted[(int64_t)e * 4 + k]];
      if (fv < 0) continue;
      for (auto &kv : P[fv]) crow[d][kv.first] += gv * kv.second;
    }
  }
  for (int64_t d = 0; d < N; d++) {
    for (auto &kv : crow[d])
      if (kv.second != 0.0) {
        c.row_col.push_back(kv.first);
        c.row_val.push_back(kv.second);
      }
    c.row_ptr[d + 1] = (int32_t)c.row_col.size();
  }
  // transpose
  c.col_ptr.assign(c.nc + 1, 0);
  for (int32_t a : c.row_col) c.col_ptr[a + 1]++;
  for (int a = 0; a < c.nc; a++) c.col_ptr[a + 1] += c.col_ptr[a];
  c.col_row.resize(c.row_col.size());
  c.col_val.resize(c.row_col.size());
  std::vector<int32_t> fill(c.col_ptr.begin(), c.col_ptr.end() - 1);
  for (int64_t d = 0; d < N; d++)
    for (int32_t k = c.row_ptr[d]; k < c.row_ptr[d + 1]; k++) {
      const int32_t a = c.row_col[k];
      c.col_row[fill[a]] = (int32_t)d;
      c.col_val[fill[a]] = c.row_val[k];
      fill[a]++;
    }
}

void coarse_galerkin(const Space &s, const Coarse &c, const double *vals, const int32_t *ov_nz,
                     const double *ov_val, int64_t nov, std::vector<double> &Ac, int nthreads) {
  const int nc = c.nc;
  Ac.assign((size_t)nc * nc, 0.0);
  if (nc == 0) return;
  auto add_entry = [&](std::vector<double> &acc, int32_t i, int32_t j, double a) {
    for (int32_t ki = c.row_ptr[i]; ki < c.row_ptr[i + 1]; ki++) {
      const double ci = c.row_val[ki] * a;
      double *row = &acc[(size_t)c.row_col[ki] * nc];
      for (int32_t kj = c.row_ptr[j]; kj < c.row_ptr[j + 1]; kj++) row[c.row_col[kj]] += ci * c.row_val[kj];
    }
  };
  if (vals) {
    const int64_t N = s.ndof;
    nthreads = std::max(1, nthreads);
    std::vector<std::vector<double>> part(nthreads, std::vector<double>((size_t)nc * nc, 0.0));
    std::vector<std::thread> th;
    const int64_t chunk = (N + nthreads - 1) / nthreads;
    for (int t = 0; t < nthreads; t++)
      th.emplace_back([&, t]() {
        const int64_t lo = t * chunk, hi = std::min(N, lo + chunk);
        for (int64_t i = lo; i < hi; i++) {
          if (c.row_ptr[i] == c.row_ptr[i + 1]) continue;
          for (int32_t nz = s.rowptr[i]; nz < s.rowptr[i + 1]; nz++)
            if (vals[nz] != 0.0) add_entry(part[t], (int32_t)i, s.colind[nz], vals[nz]);
        }
      });
    for (auto &t : th) t.join();
    for (auto &p : part)
      for (size_t k = 0; k < Ac.size(); k++) Ac[k] += p[k];
  } else {
    for (int64_t q = 0; q < nov; q++) {
      const int32_t nz = ov_nz[q];
      const int32_t i = (int32_t)(std::upper_bound(s.rowptr.begin(), s.rowptr.end(), nz) - s.rowptr.begin() - 1);
      add_entry(Ac, i, s.colind[nz], ov_val[q]);
    }
  }
}

}  // namespace opem
