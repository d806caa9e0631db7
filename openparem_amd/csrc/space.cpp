// Host-side topology and DOF numbering (see space.h).
//
// Numbering convention: every tetrahedron's vertices are taken in increasing global id;
// entities are visited element by element in input order, 6 edges -> 4 faces -> cell per
// element, and an entity gets its k consecutive DOFs the first time it is visited.  DOFs
// inside an entity follow the basis pattern order of nd_tables.cpp.
#include "space.h"

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <thread>
#include <unordered_map>

#include "nd_tables.h"

namespace opem {
namespace {

const int TET_E[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
const int TET_F[4][3] = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}};
const int TRI_E[3][2] = {{0, 1}, {0, 2}, {1, 2}};

struct FaceKey {
  int32_t a, b, c;
  bool operator==(const FaceKey &o) const { return a == o.a && b == o.b && c == o.c; }
};
struct FaceHash {
  size_t operator()(const FaceKey &k) const {
    uint64_t h = (uint64_t)(uint32_t)k.a * 0x9E3779B97F4A7C15ull;
    h ^= (uint64_t)(uint32_t)k.b * 0xC2B2AE3D27D4EB4Full + (h << 6) + (h >> 2);
    h ^= (uint64_t)(uint32_t)k.c * 0x165667B19E3779F9ull + (h << 6) + (h >> 2);
    return (size_t)h;
  }
};

template <class F>
void parallel_for(int64_t n, int nthreads, F f) {
  if (nthreads <= 1 || n < 4096) {
    f(0, n);
    return;
  }
  std::vector<std::thread> th;
  int64_t chunk = (n + nthreads - 1) / nthreads;
  for (int t = 0; t < nthreads; t++) {
    int64_t lo = t * chunk, hi = std::min(n, lo + chunk);
    if (lo >= hi) break;
    th.emplace_back(f, lo, hi);
  }
  for (auto &t : th) t.join();
}

}  // namespace

void build_space(const Mesh &m, int p, Space &s, int nthreads) {
  if (p < 1 || p > 3) throw std::runtime_error("order must be 1, 2 or 3");
  if (m.nt <= 0 || m.nv <= 0) throw std::runtime_error("empty mesh");
  s = Space();
  s.p = p;
  s.nd = nd_tet_ndof(p);
  s.ndf = nd_tri_ndof(p);
  EntityCounts ec = nd_entity_counts(p);
  s.k_edge = ec.edge;
  s.k_face = ec.face;
  s.k_cell = ec.cell;
  const int64_t nt = m.nt, nv = m.nv;
  s.tets_sorted.resize(nt * 4);
  for (int64_t e = 0; e < nt; e++) {
    int32_t v[4];
    for (int k = 0; k < 4; k++) {
      v[k] = m.tets[e * 4 + k];
      if (v[k] < 0 || v[k] >= nv) throw std::runtime_error("tet vertex index out of range");
    }
    std::sort(v, v + 4);
    if (v[0] == v[1] || v[1] == v[2] || v[2] == v[3]) throw std::runtime_error("degenerate tet");
    for (int k = 0; k < 4; k++) s.tets_sorted[e * 4 + k] = v[k];
  }
  // entity numbering by first visit
  std::unordered_map<int64_t, int64_t> edge_off;
  std::unordered_map<FaceKey, int64_t, FaceHash> face_off;
  edge_off.reserve(nt * 2);
  face_off.reserve(nt * 3);
  const int nd = s.nd;
  s.elem_dofs.resize(nt * nd);
  int64_t next = 0;
  for (int64_t e = 0; e < nt; e++) {
    const int32_t *v = &s.tets_sorted[e * 4];
    int col = 0;
    for (int k = 0; k < 6; k++) {
      int64_t key = (int64_t)v[TET_E[k][0]] * nv + v[TET_E[k][1]];
      auto it = edge_off.find(key);
      int64_t off;
      if (it == edge_off.end()) {
        off = next;
        edge_off.emplace(key, off);
        next += ec.edge;
      } else {
        off = it->second;
      }
      for (int t = 0; t < ec.edge; t++) s.elem_dofs[e * nd + col++] = (int32_t)(off + t);
    }
    for (int k = 0; k < 4; k++) {
      FaceKey key{v[TET_F[k][0]], v[TET_F[k][1]], v[TET_F[k][2]]};
      int64_t off = -1;
      if (ec.face) {
        auto it = face_off.find(key);
        if (it == face_off.end()) {
          off = next;
          face_off.emplace(key, off);
          next += ec.face;
        } else {
          off = it->second;
        }
      } else {
        face_off.emplace(key, -1);
      }
      for (int t = 0; t < ec.face; t++) s.elem_dofs[e * nd + col++] = (int32_t)(off + t);
    }
    for (int t = 0; t < ec.cell; t++) s.elem_dofs[e * nd + col++] = (int32_t)(next + t);
    next += ec.cell;
    if (next >= (int64_t)INT32_MAX) throw std::runtime_error("DOF count exceeds int32");
  }
  s.ndof = next;
  // boundary faces
  const int ndf = s.ndf;
  s.bf_sorted.resize(m.nbf * 3);
  s.bface_dofs.resize(m.nbf * ndf);
  for (int64_t f = 0; f < m.nbf; f++) {
    int32_t v[3] = {m.bfaces[f * 3], m.bfaces[f * 3 + 1], m.bfaces[f * 3 + 2]};
    for (int k = 0; k < 3; k++)
      if (v[k] < 0 || v[k] >= nv) throw std::runtime_error("boundary face vertex out of range");
    std::sort(v, v + 3);
    for (int k = 0; k < 3; k++) s.bf_sorted[f * 3 + k] = v[k];
    int col = 0;
    for (int k = 0; k < 3; k++) {
      int64_t key = (int64_t)v[TRI_E[k][0]] * nv + v[TRI_E[k][1]];
      auto it = edge_off.find(key);
      if (it == edge_off.end()) throw std::runtime_error("boundary face edge not in mesh");
      for (int t = 0; t < ec.edge; t++) s.bface_dofs[f * ndf + col++] = (int32_t)(it->second + t);
    }
    auto it = face_off.find(FaceKey{v[0], v[1], v[2]});
    if (it == face_off.end()) throw std::runtime_error("boundary face not in mesh");
    for (int t = 0; t < ec.face; t++) s.bface_dofs[f * ndf + col++] = (int32_t)(it->second + t);
  }
  // DOF -> element occurrences
  const int64_t N = s.ndof;
  s.d2e_ptr.assign(N + 1, 0);
  for (int64_t i = 0; i < nt * nd; i++) s.d2e_ptr[s.elem_dofs[i] + 1]++;
  for (int64_t i = 0; i < N; i++) s.d2e_ptr[i + 1] += s.d2e_ptr[i];
  s.d2e.resize(nt * nd);
  {
    std::vector<int32_t> fill(s.d2e_ptr.begin(), s.d2e_ptr.end() - 1);
    for (int64_t e = 0; e < nt; e++)
      for (int a = 0; a < nd; a++) s.d2e[fill[s.elem_dofs[e * nd + a]]++] = (int32_t)(e * nd + a);
  }
  s.owner_elem.resize(N);
  s.owner_local.resize(N);
  for (int64_t r = 0; r < N; r++) {
    int32_t pk = s.d2e[s.d2e_ptr[r]];
    s.owner_elem[r] = pk / nd;
    s.owner_local[r] = pk % nd;
  }
  // owner of each boundary face: first element containing all its DOF entities
  s.bface_owner.resize(m.nbf);
  for (int64_t f = 0; f < m.nbf; f++) {
    // the face DOF of ND1 does not exist; identify by vertices instead
    const int32_t *v = &s.bf_sorted[f * 3];
    int32_t d0 = s.bface_dofs[f * ndf];  // first edge DOF: candidate elements
    int32_t own = -1;
    for (int32_t k = s.d2e_ptr[d0]; k < s.d2e_ptr[d0 + 1] && own < 0; k++) {
      int32_t e = s.d2e[k] / nd;
      const int32_t *tv = &s.tets_sorted[e * 4];
      int hit = 0;
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 4; b++)
          if (v[a] == tv[b]) hit++;
      if (hit == 3) own = e;
    }
    if (own < 0) throw std::runtime_error("boundary face has no owning tet");
    s.bface_owner[f] = own;
  }
  // CSR pattern: row r = union of DOFs of the elements containing r
  std::vector<int32_t> rowlen(N, 0);
  auto row_cols = [&](int64_t r, std::vector<int32_t> &buf) {
    buf.clear();
    for (int32_t k = s.d2e_ptr[r]; k < s.d2e_ptr[r + 1]; k++) {
      int32_t e = s.d2e[k] / nd;
      for (int b = 0; b < nd; b++) buf.push_back(s.elem_dofs[(int64_t)e * nd + b]);
    }
    std::sort(buf.begin(), buf.end());
    buf.erase(std::unique(buf.begin(), buf.end()), buf.end());
  };
  parallel_for(N, nthreads, [&](int64_t lo, int64_t hi) {
    std::vector<int32_t> buf;
    for (int64_t r = lo; r < hi; r++) {
      row_cols(r, buf);
      rowlen[r] = (int32_t)buf.size();
    }
  });
  s.rowptr.assign(N + 1, 0);
  int64_t acc = 0;
  for (int64_t r = 0; r < N; r++) {
    s.rowptr[r] = (int32_t)acc;
    acc += rowlen[r];
    if (acc >= (int64_t)INT32_MAX) throw std::runtime_error("nnz exceeds int32");
  }
  s.rowptr[N] = (int32_t)acc;
  s.nnz = acc;
  s.colind.resize(acc);
  parallel_for(N, nthreads, [&](int64_t lo, int64_t hi) {
    std::vector<int32_t> buf;
    for (int64_t r = lo; r < hi; r++) {
      row_cols(r, buf);
      std::copy(buf.begin(), buf.end(), s.colind.begin() + s.rowptr[r]);
    }
  });
  // geometric factors
  s.gm.resize(nt * 6);
  s.gk.resize(nt * 6);
  s.detj.resize(nt);
  parallel_for(nt, nthreads, [&](int64_t lo, int64_t hi) {
    for (int64_t e = lo; e < hi; e++) {
      double v[4][3];
      for (int k = 0; k < 4; k++)
        for (int d = 0; d < 3; d++) v[k][d] = m.xyz[(int64_t)s.tets_sorted[e * 4 + k] * 3 + d];
      s.detj[e] = tet_factors(v, &s.gm[e * 6], &s.gk[e * 6]);
    }
  });
  for (int64_t e = 0; e < nt; e++)
    if (!(std::fabs(s.detj[e]) > 0)) throw std::runtime_error("zero-volume tet");
}

void build_patches(const Space &s, const std::vector<uint8_t> &exclude, Patches &P, int nthreads) {
  P = Patches();
  const int nd = s.nd;
  const int64_t nt = (int64_t)s.tets_sorted.size() / 4;
  // entity vertex sets per element-local DOF
  std::vector<std::vector<int>> ent(nd);
  {
    int a = 0;
    for (int k = 0; k < 6; k++)
      for (int t = 0; t < s.k_edge; t++) ent[a++] = {TET_E[k][0], TET_E[k][1]};
    for (int k = 0; k < 4; k++)
      for (int t = 0; t < s.k_face; t++) ent[a++] = {TET_F[k][0], TET_F[k][1], TET_F[k][2]};
    for (int t = 0; t < s.k_cell; t++) ent[a++] = {0, 1, 2, 3};
  }
  int32_t nv = 0;
  for (int32_t v : s.tets_sorted) nv = std::max(nv, v + 1);
  std::vector<std::vector<int32_t>> pv(nv);
  for (int64_t e = 0; e < nt; e++)
    for (int a = 0; a < nd; a++) {
      const int32_t d = s.elem_dofs[e * nd + a];
      if (exclude[d]) continue;
      for (int k : ent[a]) pv[s.tets_sorted[e * 4 + k]].push_back(d);
    }
  P.idx_off.push_back(0);
  P.val_off.push_back(0);
  for (int32_t v = 0; v < nv; v++) {
    auto &L = pv[v];
    std::sort(L.begin(), L.end());
    L.erase(std::unique(L.begin(), L.end()), L.end());
    if (L.empty()) continue;
    P.idx.insert(P.idx.end(), L.begin(), L.end());
    P.idx_off.push_back((int32_t)P.idx.size());
    P.val_off.push_back(P.val_off.back() + (int64_t)L.size() * (int64_t)L.size());
    P.nmax = std::max(P.nmax, (int)L.size());
    std::vector<int32_t>().swap(L);
  }
  P.npatch = (int64_t)P.idx_off.size() - 1;
  P.nz.resize(P.val_off.back());
  parallel_for(P.npatch, nthreads, [&](int64_t lo, int64_t hi) {
    for (int64_t p = lo; p < hi; p++) {
      const int32_t *I = &P.idx[P.idx_off[p]];
      const int n = P.idx_off[p + 1] - P.idx_off[p];
      for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) P.nz[P.val_off[p] + (int64_t)i * n + j] = (int32_t)csr_find(s, I[i], I[j]);
    }
  });
  P.d2p_ptr.assign(s.ndof + 1, 0);
  for (int32_t d : P.idx) P.d2p_ptr[d + 1]++;
  for (int64_t i = 0; i < s.ndof; i++) P.d2p_ptr[i + 1] += P.d2p_ptr[i];
  P.d2p_y.resize(P.idx.size());
  std::vector<int32_t> fill(P.d2p_ptr.begin(), P.d2p_ptr.end() - 1);
  for (int32_t k = 0; k < (int32_t)P.idx.size(); k++) P.d2p_y[fill[P.idx[k]]++] = k;
}

int64_t csr_find(const Space &s, int32_t row, int32_t col) {
  auto b = s.colind.begin() + s.rowptr[row], e = s.colind.begin() + s.rowptr[row + 1];
  auto it = std::lower_bound(b, e, col);
  if (it == e || *it != col) return -1;
  return it - s.colind.begin();
}

void face_matrices(const Mesh &m, const Space &s, int64_t f, double *Kf, double *Mf) {
  const TriTables &T = tri_tables(s.p);
  double v[3][3];
  for (int k = 0; k < 3; k++)
    for (int d = 0; d < 3; d++) v[k][d] = m.xyz[(int64_t)s.bf_sorted[f * 3 + k] * 3 + d];
  double fm[3], fk;
  tri_factors(v, fm, &fk);
  const int n = T.nd;
  for (int i = 0; i < n * n; i++) {
    Mf[i] = fm[0] * T.M[i] + fm[1] * T.M[n * n + i] + fm[2] * T.M[2 * n * n + i];
    Kf[i] = fk * T.K[i];
  }
}

}  // namespace opem
