// Conforming local refinement of a tetrahedral mesh by edge bisection (the step after the marking of
// fem3D::calculateMeshErrors; the reference calls MFEM's ParMesh::GeneralRefinement, fem3D.cpp:1384-1447).
//
// Each marked tet that has not already been split is bisected at the midpoint of its longest edge (ties: the
// smaller vertex pair).  Bisecting an edge splits every tet and every boundary triangle that contains it, so the
// mesh stays conforming after each step: a tet (a, b, c, d) split at edge (a, b) becomes (m, b, c, d) and
// (a, m, c, d), which keeps its orientation; children keep the parent's attribute.  Host only.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <unordered_map>
#include <vector>

namespace {

inline int64_t ekey(int32_t a, int32_t b) {
  return a < b ? ((int64_t)a << 32 | (uint32_t)b) : ((int64_t)b << 32 | (uint32_t)a);
}

struct Refiner {
  std::vector<double> xyz;
  std::vector<int32_t> tets, tattr, tpar, bf, battr;   // tpar: the input tet each tet descends from
  std::vector<uint8_t> talive, balive;
  std::unordered_map<int64_t, std::vector<int32_t>> estar, bstar;   // edge -> tets / boundary faces holding it

  void add_tet(int32_t t) {
    const int32_t *v = &tets[4 * (size_t)t];
    for (int i = 0; i < 4; i++)
      for (int j = i + 1; j < 4; j++) estar[ekey(v[i], v[j])].push_back(t);
  }
  void drop(std::unordered_map<int64_t, std::vector<int32_t>> &m, int64_t k, int32_t id) {
    auto &s = m[k];
    s.erase(std::find(s.begin(), s.end(), id));
  }
  void add_bf(int32_t f) {
    const int32_t *v = &bf[3 * (size_t)f];
    for (int i = 0; i < 3; i++)
      for (int j = i + 1; j < 3; j++) bstar[ekey(v[i], v[j])].push_back(f);
  }

  void bisect(int32_t a, int32_t b) {
    const int32_t m = (int32_t)(xyz.size() / 3);
    for (int d = 0; d < 3; d++) xyz.push_back(0.5 * (xyz[3 * (size_t)a + d] + xyz[3 * (size_t)b + d]));
    const int64_t k = ekey(a, b);
    const std::vector<int32_t> ts = estar[k];
    for (int32_t t : ts) {
      int32_t v[4];
      std::memcpy(v, &tets[4 * (size_t)t], 16);
      for (int i = 0; i < 4; i++)
        for (int j = i + 1; j < 4; j++) drop(estar, ekey(v[i], v[j]), t);
      talive[t] = 0;
      for (int side = 0; side < 2; side++) {   // side 0 replaces a by m, side 1 replaces b by m
        int32_t w[4];
        for (int i = 0; i < 4; i++) w[i] = v[i] == (side == 0 ? a : b) ? m : v[i];
        const int32_t id = (int32_t)(tets.size() / 4);
        tets.insert(tets.end(), w, w + 4);
        tattr.push_back(tattr[t]);
        tpar.push_back(tpar[t]);
        talive.push_back(1);
        add_tet(id);
      }
    }
    const std::vector<int32_t> fs = bstar[k];
    for (int32_t f : fs) {
      int32_t v[3];
      std::memcpy(v, &bf[3 * (size_t)f], 12);
      for (int i = 0; i < 3; i++)
        for (int j = i + 1; j < 3; j++) drop(bstar, ekey(v[i], v[j]), f);
      balive[f] = 0;
      for (int side = 0; side < 2; side++) {
        int32_t w[3];
        for (int i = 0; i < 3; i++) w[i] = v[i] == (side == 0 ? a : b) ? m : v[i];
        const int32_t id = (int32_t)(bf.size() / 3);
        bf.insert(bf.end(), w, w + 3);
        battr.push_back(battr[f]);
        balive.push_back(1);
        add_bf(id);
      }
    }
    estar.erase(k);
    bstar.erase(k);
  }

  void longest_edge(int32_t t, int32_t &ea, int32_t &eb) const {
    const int32_t *v = &tets[4 * (size_t)t];
    double best = -1;
    int64_t bk = 0;
    for (int i = 0; i < 4; i++)
      for (int j = i + 1; j < 4; j++) {
        double l2 = 0;
        for (int d = 0; d < 3; d++) {
          const double x = xyz[3 * (size_t)v[i] + d] - xyz[3 * (size_t)v[j] + d];
          l2 += x * x;
        }
        const int64_t k = ekey(v[i], v[j]);
        if (l2 > best || (l2 == best && k < bk)) {
          best = l2;
          bk = k;
          ea = v[i];
          eb = v[j];
        }
      }
  }
};

}  // namespace

extern "C" {

// Refines the marked tets (0-based ids into tets).  Outputs: the refined mesh's vertices (the input vertices
// first, in order), alive tets and boundary faces in creation order, with attributes; parent (optional): the
// input tet each output tet descends from.  Call with xyz_out == NULL to get the counts *nv_out, *nt_out,
// *nbf_out, then again with arrays of those sizes.  Returns 0, or -2 for bad input.
int opem_refine_mesh(int64_t nv, const double *xyz, int64_t nt, const int32_t *tets, const int32_t *tet_attr,
                     int64_t nbf, const int32_t *bfaces, const int32_t *bface_attr, int64_t nmarked,
                     const int32_t *marked, int64_t *nv_out, double *xyz_out, int64_t *nt_out, int32_t *tets_out,
                     int32_t *tet_attr_out, int64_t *nbf_out, int32_t *bfaces_out, int32_t *bface_attr_out,
                     int32_t *parent) {
  if (nv < 4 || nt < 1 || nbf < 0 || nmarked < 0 || !xyz || !tets || !tet_attr || (nbf > 0 && (!bfaces || !bface_attr)) ||
      (nmarked > 0 && !marked) || !nv_out || !nt_out || !nbf_out)
    return -2;
  if (nv > INT32_MAX / 2 || nt > INT32_MAX / 16) return -2;
  for (int64_t i = 0; i < 4 * nt; i++)
    if (tets[i] < 0 || tets[i] >= nv) return -2;
  for (int64_t i = 0; i < 3 * nbf; i++)
    if (bfaces[i] < 0 || bfaces[i] >= nv) return -2;
  for (int64_t i = 0; i < nmarked; i++)
    if (marked[i] < 0 || marked[i] >= nt) return -2;
  Refiner R;
  R.xyz.assign(xyz, xyz + 3 * nv);
  R.tets.assign(tets, tets + 4 * nt);
  R.tattr.assign(tet_attr, tet_attr + nt);
  R.talive.assign(nt, 1);
  R.tpar.resize(nt);
  for (int64_t t = 0; t < nt; t++) R.tpar[t] = (int32_t)t;
  if (nbf > 0) {
    R.bf.assign(bfaces, bfaces + 3 * nbf);
    R.battr.assign(bface_attr, bface_attr + nbf);
  }
  R.balive.assign(nbf, 1);
  for (int32_t t = 0; t < (int32_t)nt; t++) R.add_tet(t);
  for (int32_t f = 0; f < (int32_t)nbf; f++) R.add_bf(f);
  for (int64_t i = 0; i < nmarked; i++) {
    const int32_t t = marked[i];
    if (!R.talive[t]) continue;   // already split by a neighbour's bisection
    int32_t a = 0, b = 0;
    R.longest_edge(t, a, b);
    R.bisect(a, b);
  }
  int64_t cnt_t = 0, cnt_f = 0;
  for (size_t t = 0; t < R.talive.size(); t++) cnt_t += R.talive[t];
  for (size_t f = 0; f < R.balive.size(); f++) cnt_f += R.balive[f];
  *nv_out = (int64_t)R.xyz.size() / 3;
  *nt_out = cnt_t;
  *nbf_out = cnt_f;
  if (!xyz_out) return 0;
  if (!tets_out || !tet_attr_out || (cnt_f > 0 && (!bfaces_out || !bface_attr_out))) return -2;
  std::memcpy(xyz_out, R.xyz.data(), R.xyz.size() * 8);
  int64_t o = 0;
  for (size_t t = 0; t < R.talive.size(); t++)
    if (R.talive[t]) {
      std::memcpy(tets_out + 4 * o, &R.tets[4 * t], 16);
      tet_attr_out[o] = R.tattr[t];
      if (parent) parent[o] = R.tpar[t];
      o++;
    }
  o = 0;
  for (size_t f = 0; f < R.balive.size(); f++)
    if (R.balive[f]) {
      std::memcpy(bfaces_out + 3 * o, &R.bf[3 * f], 12);
      bface_attr_out[o] = R.battr[f];
      o++;
    }
  return 0;
}

}  // extern "C"
