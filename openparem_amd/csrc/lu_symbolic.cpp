// Symbolic phase of the multifrontal sparse LU (see lu_symbolic.h).
#include "lu_symbolic.h"

#include <algorithm>
#include <cmath>
#include <stdexcept>

namespace opem {
namespace {

const int TET_E[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
const int TET_F[4][3] = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}};

// Nested-dissection tree over the vertices of a graph: node parents (preorder of creation) and each vertex's
// node.
struct NdTree {
  std::vector<int32_t> parent, depth;
  std::vector<int32_t> vnode;
  int add(int par) {
    parent.push_back(par);
    depth.push_back(par < 0 ? 0 : depth[par] + 1);
    return (int)parent.size() - 1;
  }
};

// Recursive bisection with an explicit stack.  split(V, id, A, B, S) partitions V into A, B and a separator S
// with no edge between A and B; it returns false when it cannot (the set becomes a leaf).
template <class Split>
void dissect(int64_t nv, int leaf, Split split, NdTree &nt) {
  nt.vnode.assign(nv, -1);
  struct Task {
    std::vector<int32_t> v;
    int parent;
  };
  std::vector<Task> stack;
  {
    Task t;
    t.v.resize(nv);
    for (int64_t i = 0; i < nv; i++) t.v[i] = (int32_t)i;
    t.parent = -1;
    stack.push_back(std::move(t));
  }
  int id = 0;
  std::vector<int32_t> A, B, S;
  while (!stack.empty()) {
    Task t = std::move(stack.back());
    stack.pop_back();
    if (t.v.empty()) continue;
    A.clear();
    B.clear();
    S.clear();
    if ((int64_t)t.v.size() <= leaf || !split(t.v, id++, A, B, S) || (A.empty() && B.empty())) {
      const int node = nt.add(t.parent);
      for (int32_t v : t.v) nt.vnode[v] = node;
      continue;
    }
    const int node = nt.add(t.parent);
    for (int32_t v : S) nt.vnode[v] = node;
    stack.push_back(Task{B, node});
    stack.push_back(Task{A, node});
  }
}

// Symmetric adjacency (no self loops) of a CSR pattern restricted to rows/cols with keep[] != 0.
void sym_adjacency(int64_t n, const int32_t *rp, const int32_t *ci, const std::vector<uint8_t> &keep,
                   std::vector<int64_t> &ap, std::vector<int32_t> &ac) {
  std::vector<int64_t> cnt(n + 1, 0);
  for (int64_t i = 0; i < n; i++) {
    if (!keep[i]) continue;
    for (int32_t k = rp[i]; k < rp[i + 1]; k++) {
      const int32_t j = ci[k];
      if (j == i || !keep[j]) continue;
      cnt[i + 1]++;
      cnt[j + 1]++;
    }
  }
  for (int64_t i = 0; i < n; i++) cnt[i + 1] += cnt[i];
  std::vector<int32_t> tmp(cnt[n]);
  std::vector<int64_t> fill(cnt.begin(), cnt.end() - 1);
  for (int64_t i = 0; i < n; i++) {
    if (!keep[i]) continue;
    for (int32_t k = rp[i]; k < rp[i + 1]; k++) {
      const int32_t j = ci[k];
      if (j == i || !keep[j]) continue;
      tmp[fill[i]++] = j;
      tmp[fill[j]++] = (int32_t)i;
    }
  }
  ap.assign(n + 1, 0);
  ac.clear();
  ac.reserve(tmp.size());
  for (int64_t i = 0; i < n; i++) {
    auto b = tmp.begin() + cnt[i], e = tmp.begin() + cnt[i + 1];
    std::sort(b, e);
    auto u = std::unique(b, e);
    ac.insert(ac.end(), b, u);
    ap[i + 1] = (int64_t)ac.size();
  }
}

// Smallest vertex separator inside the boundary of a cut: SA (side-a vertices with a side-b neighbour) and SB
// (the converse) with the cut edges between them form a bipartite graph, and any vertex cover of it separates the
// two sides.  A maximum matching (augmenting paths) gives a minimum cover by Koenig's theorem: the SA vertices
// not reachable from an unmatched SA vertex by alternating paths, plus the SB vertices that are.
void min_cover(const std::vector<int32_t> &SA, const std::vector<int32_t> &SB, const std::vector<int64_t> &vp,
               const std::vector<int32_t> &vc, const std::vector<int32_t> &mark, int side_a, int side_b,
               std::vector<int32_t> &cover) {
  cover.clear();
  const int na = (int)SA.size(), nb = (int)SB.size();
  if (na == 0 || nb == 0) return;
  std::vector<int32_t> all(SB);
  std::sort(all.begin(), all.end());
  auto bidx = [&](int32_t v) -> int {
    auto it = std::lower_bound(all.begin(), all.end(), v);
    return (it != all.end() && *it == v) ? (int)(it - all.begin()) : -1;
  };
  // bipartite adjacency a -> b (b indexed in the sorted SB)
  std::vector<int32_t> ap(na + 1, 0), ad;
  for (int i = 0; i < na; i++) {
    const int32_t v = SA[i];
    for (int64_t q = vp[v]; q < vp[v + 1]; q++)
      if (mark[vc[q]] == side_b) {
        const int j = bidx(vc[q]);
        if (j >= 0) ad.push_back(j);
      }
    ap[i + 1] = (int32_t)ad.size();
  }
  (void)side_a;
  std::vector<int32_t> ma(na, -1), mb(nb, -1), vis(nb, -1);
  // augmenting paths by iterative DFS (Kuhn), greedy start
  for (int i = 0; i < na; i++)
    for (int32_t k = ap[i]; k < ap[i + 1]; k++)
      if (mb[ad[k]] < 0) {
        ma[i] = ad[k];
        mb[ad[k]] = i;
        break;
      }
  std::vector<std::pair<int, int32_t>> st;
  std::vector<int> from(nb);
  for (int i = 0; i < na; i++) {
    if (ma[i] >= 0) continue;
    st.clear();
    st.push_back({i, ap[i]});
    bool found = false;
    while (!st.empty() && !found) {
      auto &top = st.back();
      const int a = top.first;
      if (top.second >= ap[a + 1]) {
        st.pop_back();
        continue;
      }
      const int b = ad[top.second++];
      if (vis[b] == i) continue;
      vis[b] = i;
      if (mb[b] < 0) {
        // augment along the stack
        int bb = b;
        for (int s = (int)st.size() - 1; s >= 0; s--) {
          const int aa = st[s].first;
          const int prev = ma[aa];
          ma[aa] = bb;
          mb[bb] = aa;
          bb = prev;
        }
        found = true;
      } else {
        st.push_back({mb[b], ap[mb[b]]});
      }
    }
  }
  // Koenig: alternating reachability from unmatched a-vertices
  std::vector<uint8_t> za(na, 0), zb(nb, 0);
  std::vector<int> q;
  for (int i = 0; i < na; i++)
    if (ma[i] < 0) {
      za[i] = 1;
      q.push_back(i);
    }
  for (size_t h = 0; h < q.size(); h++) {
    const int a = q[h];
    for (int32_t k = ap[a]; k < ap[a + 1]; k++) {
      const int b = ad[k];
      if (zb[b] || ma[a] == b) continue;  // only non-matching edges from the a side
      zb[b] = 1;
      const int a2 = mb[b];
      if (a2 >= 0 && !za[a2]) {
        za[a2] = 1;
        q.push_back(a2);
      }
    }
  }
  for (int i = 0; i < na; i++)
    if (!za[i]) cover.push_back(SA[i]);
  for (int j = 0; j < nb; j++)
    if (zb[j]) cover.push_back(all[j]);
}

// From a DOF -> node assignment over a tree (parents by node id, -1 for roots; dof_node -1 = excluded) to the
// full LuTree: empty nodes are spliced out, nodes are put in postorder, DOFs numbered node by node, update sets
// computed bottom-up from the symmetric adjacency (ap, ac), then the extend-add and assembly maps over the
// caller's pattern (rp, ci).
void finish_tree(int64_t ndof, const std::vector<int64_t> &ap, const std::vector<int32_t> &ac, const int32_t *rp,
                 const int32_t *ci, const std::vector<int32_t> &tparent, const std::vector<int32_t> &dof_node,
                 int last_tnode, LuTree &T) {
  const int nt = (int)tparent.size();
  std::vector<int64_t> cnt(nt, 0);
  for (int64_t d = 0; d < ndof; d++)
    if (dof_node[d] >= 0) cnt[dof_node[d]]++;
  // effective parent: nearest non-empty ancestor
  std::vector<int32_t> eff(nt);
  for (int t = 0; t < nt; t++) {
    int p = tparent[t];
    while (p >= 0 && cnt[p] == 0) p = tparent[p];
    eff[t] = p;
  }
  std::vector<std::vector<int32_t>> kids(nt);
  std::vector<int32_t> roots;
  for (int t = 0; t < nt; t++) {
    if (cnt[t] == 0) continue;
    if (eff[t] < 0)
      roots.push_back(t);
    else
      kids[eff[t]].push_back(t);
  }
  // postorder (iterative DFS)
  std::vector<int32_t> post;
  post.reserve(nt);
  {
    std::vector<std::pair<int32_t, size_t>> st;
    for (int32_t r : roots) {
      st.push_back({r, 0});
      while (!st.empty()) {
        auto &top = st.back();
        if (top.second < kids[top.first].size()) {
          const int32_t c = kids[top.first][top.second++];
          st.push_back({c, 0});
        } else {
          post.push_back(top.first);
          st.pop_back();
        }
      }
    }
  }
  const int nn = (int)post.size();
  std::vector<int32_t> newid(nt, -1);
  for (int i = 0; i < nn; i++) newid[post[i]] = i;
  T = LuTree();
  T.ndof = ndof;
  T.nnodes = nn;
  T.parent.assign(nn, -1);
  T.last_node = last_tnode >= 0 && cnt[last_tnode] > 0 ? newid[last_tnode] : -1;
  for (int i = 0; i < nn; i++) T.parent[i] = eff[post[i]] >= 0 ? newid[eff[post[i]]] : -1;
  // DOF numbering: nodes in postorder, DOFs ascending inside a node
  std::vector<std::vector<int32_t>> ndofs(nn);
  for (int64_t d = 0; d < ndof; d++)
    if (dof_node[d] >= 0) ndofs[newid[dof_node[d]]].push_back((int32_t)d);
  T.dbeg.assign(nn + 1, 0);
  T.iperm.assign(ndof, -1);
  for (int i = 0; i < nn; i++) {
    T.dbeg[i + 1] = T.dbeg[i] + (int64_t)ndofs[i].size();
    for (size_t k = 0; k < ndofs[i].size(); k++) {
      T.iperm[ndofs[i][k]] = (int32_t)(T.dbeg[i] + (int64_t)k);
      T.perm.push_back(ndofs[i][k]);
    }
  }
  T.n = T.dbeg[nn];
  // update sets, bottom-up (children precede parents in postorder)
  std::vector<std::vector<int32_t>> kid_new(nn);
  for (int i = 0; i < nn; i++)
    if (T.parent[i] >= 0) kid_new[T.parent[i]].push_back(i);
  T.uoff.assign(nn + 1, 0);
  std::vector<std::vector<int32_t>> U(nn);
  std::vector<int32_t> buf;
  for (int t = 0; t < nn; t++) {
    const int64_t dend = T.dbeg[t + 1];
    buf.clear();
    for (int64_t pos = T.dbeg[t]; pos < dend; pos++) {
      const int32_t d = T.perm[pos];
      for (int64_t k = ap[d]; k < ap[d + 1]; k++) {
        const int32_t g = T.iperm[ac[k]];
        if (g >= dend) buf.push_back(g);
      }
    }
    for (int32_t c : kid_new[t])
      for (int32_t g : U[c])
        if (g >= dend) buf.push_back(g);
    std::sort(buf.begin(), buf.end());
    buf.erase(std::unique(buf.begin(), buf.end()), buf.end());
    U[t] = buf;
    T.uoff[t + 1] = T.uoff[t] + (int64_t)buf.size();
  }
  T.uidx.resize(T.uoff[nn]);
  for (int t = 0; t < nn; t++) std::copy(U[t].begin(), U[t].end(), T.uidx.begin() + T.uoff[t]);
  // levels, slots, front offsets, flops
  T.level.assign(nn, 0);
  T.slot.assign(nn, 0);
  for (int t = 0; t < nn; t++) {
    int lv = 0;
    for (size_t k = 0; k < kid_new[t].size(); k++) {
      lv = std::max(lv, T.level[kid_new[t][k]] + 1);
      T.slot[kid_new[t][k]] = (int32_t)k;
    }
    T.level[t] = lv;
    T.nlevels = std::max(T.nlevels, lv + 1);
    T.max_children = std::max(T.max_children, (int)kid_new[t].size());
  }
  T.foff.assign(nn + 1, 0);
  for (int t = 0; t < nn; t++) {
    const double a = (double)T.nd(t), b = (double)T.nu(t);
    T.foff[t + 1] = T.foff[t] + T.nf(t) * T.nf(t);
    T.max_front = std::max(T.max_front, T.nf(t));
    T.flops += 8.0 * (a * a * a / 3.0 + a * a * b + a * b * b);
  }
  // extend-add map: every update entry of a child lies in its parent's front
  T.umap.resize(T.uidx.size());
  for (int c = 0; c < nn; c++) {
    const int p = T.parent[c];
    if (p < 0) {
      if (T.nu(c) != 0) throw std::logic_error("lu: a root with update rows");
      continue;
    }
    const int32_t *Up = T.uidx.data() + T.uoff[p];
    const int64_t nup = T.nu(p), ndp = T.nd(p);
    for (int64_t k = T.uoff[c]; k < T.uoff[c + 1]; k++) {
      const int32_t g = T.uidx[k];
      if (g >= T.dbeg[p] && g < T.dbeg[p + 1]) {
        T.umap[k] = (int32_t)(g - T.dbeg[p]);
      } else {
        const int32_t *it = std::lower_bound(Up, Up + nup, g);
        if (it == Up + nup || *it != g) throw std::logic_error("lu: update row missing from the parent front");
        T.umap[k] = (int32_t)(ndp + (it - Up));
      }
    }
  }
  // assembly map: entry (i, j) goes to the front of the node that pivots the earlier of the two
  std::vector<int32_t> node_of(T.n);
  for (int t = 0; t < nn; t++)
    for (int64_t p = T.dbeg[t]; p < T.dbeg[t + 1]; p++) node_of[p] = t;
  auto pos_in = [&](int t, int32_t g) -> int64_t {
    if (g < T.dbeg[t + 1]) return g - T.dbeg[t];
    const int32_t *Ut = T.uidx.data() + T.uoff[t];
    const int32_t *it = std::lower_bound(Ut, Ut + T.nu(t), g);
    if (it == Ut + T.nu(t) || *it != g) throw std::logic_error("lu: matrix entry outside its front");
    return T.nd(t) + (it - Ut);
  };
  const int64_t nnz = rp[ndof];
  T.asm_pos.assign(nnz, -1);
  for (int64_t i = 0; i < ndof; i++) {
    const int32_t gi = T.iperm[i];
    if (gi < 0) continue;
    for (int32_t k = rp[i]; k < rp[i + 1]; k++) {
      const int32_t gj = T.iperm[ci[k]];
      if (gj < 0) continue;
      const int t = node_of[std::min(gi, gj)];
      T.asm_pos[k] = T.foff[t] + pos_in(t, gi) + T.nf(t) * pos_in(t, gj);
    }
  }
}

}  // namespace

void lu_analyse_mesh(const Mesh &m, const Space &s, const std::vector<uint8_t> &exclude,
                     const std::vector<uint8_t> &last, int leaf_vertices, LuTree &T) {
  const int64_t nv = m.nv, nt = m.nt;
  if ((int64_t)exclude.size() != s.ndof || (int64_t)last.size() != s.ndof)
    throw std::invalid_argument("lu: exclude/last must have one entry per DOF");
  // vertex graph from the tetrahedra's edges
  std::vector<int64_t> vp(nv + 1, 0);
  std::vector<int32_t> vc;
  {
    std::vector<std::pair<int32_t, int32_t>> e;
    e.reserve(nt * 12);
    for (int64_t t = 0; t < nt; t++) {
      const int32_t *v = &s.tets_sorted[t * 4];
      for (int k = 0; k < 6; k++) {
        e.push_back({v[TET_E[k][0]], v[TET_E[k][1]]});
        e.push_back({v[TET_E[k][1]], v[TET_E[k][0]]});
      }
    }
    std::sort(e.begin(), e.end());
    e.erase(std::unique(e.begin(), e.end()), e.end());
    vc.resize(e.size());
    for (size_t k = 0; k < e.size(); k++) {
      vp[e[k].first + 1]++;
      vc[k] = e[k].second;
    }
    for (int64_t i = 0; i < nv; i++) vp[i + 1] += vp[i];
  }
  std::vector<int32_t> mark(nv, -1);
  auto split = [&](const std::vector<int32_t> &V, int id, std::vector<int32_t> &A, std::vector<int32_t> &B,
                   std::vector<int32_t> &S) -> bool {
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int32_t v : V)
      for (int d = 0; d < 3; d++) {
        lo[d] = std::min(lo[d], m.xyz[(int64_t)v * 3 + d]);
        hi[d] = std::max(hi[d], m.xyz[(int64_t)v * 3 + d]);
      }
    int ax = 0;
    for (int d = 1; d < 3; d++)
      if (hi[d] - lo[d] > hi[ax] - lo[ax]) ax = d;
    std::vector<int32_t> W(V);
    const size_t h = W.size() / 2;
    std::nth_element(W.begin(), W.begin() + h, W.end(), [&](int32_t a, int32_t b) {
      const double xa = m.xyz[(int64_t)a * 3 + ax], xb = m.xyz[(int64_t)b * 3 + ax];
      return xa < xb || (xa == xb && a < b);
    });
    for (size_t k = 0; k < W.size(); k++) mark[W[k]] = k < h ? 2 * id : 2 * id + 1;
    std::vector<int32_t> SA, SB;
    for (size_t k = 0; k < W.size(); k++) {
      const int32_t v = W[k];
      const int other = k < h ? 2 * id + 1 : 2 * id;
      bool cut = false;
      for (int64_t q = vp[v]; q < vp[v + 1] && !cut; q++) cut = mark[vc[q]] == other;
      if (cut) (k < h ? SA : SB).push_back(v);
    }
    min_cover(SA, SB, vp, vc, mark, 2 * id, 2 * id + 1, S);
    for (int32_t v : S) mark[v] = -2;  // separator
    S.clear();
    for (size_t k = 0; k < W.size(); k++) {
      const int32_t v = W[k];
      if (mark[v] == -2)
        S.push_back(v);
      else
        (k < h ? A : B).push_back(v);
    }
    for (int32_t v : W) mark[v] = -1;
    return true;
  };
  NdTree nd;
  dissect(nv, std::max(1, leaf_vertices), split, nd);
  // DOF -> deepest node among its entity's vertices
  std::vector<int32_t> dof_node(s.ndof, -1);
  std::vector<uint8_t> seen(s.ndof, 0);
  const int ndl = s.nd;
  std::vector<std::vector<int>> ent(ndl);
  {
    int a = 0;
    for (int k = 0; k < 6; k++)
      for (int t = 0; t < s.k_edge; t++) ent[a++] = {TET_E[k][0], TET_E[k][1]};
    for (int k = 0; k < 4; k++)
      for (int t = 0; t < s.k_face; t++) ent[a++] = {TET_F[k][0], TET_F[k][1], TET_F[k][2]};
    for (int t = 0; t < s.k_cell; t++) ent[a++] = {0, 1, 2, 3};
  }
  int last_node = -1;
  bool any_last = false;
  for (int64_t d = 0; d < s.ndof; d++) any_last |= last[d] && !exclude[d];
  if (any_last) {
    // the `last` root sits above every dissection root
    last_node = (int)nd.parent.size();
    for (auto &p : nd.parent)
      if (p < 0) p = last_node;
    nd.parent.push_back(-1);
    nd.depth.push_back(-1);
  }
  for (int64_t t = 0; t < nt; t++)
    for (int a = 0; a < ndl; a++) {
      const int32_t d = s.elem_dofs[t * ndl + a];
      if (seen[d]) continue;
      seen[d] = 1;
      if (exclude[d]) continue;
      if (last[d]) {
        dof_node[d] = last_node;
        continue;
      }
      int best = -1;
      for (int k : ent[a]) {
        const int nn = nd.vnode[s.tets_sorted[t * 4 + k]];
        if (best < 0 || nd.depth[nn] > nd.depth[best]) best = nn;
      }
      dof_node[d] = best;
    }
  std::vector<uint8_t> keep(s.ndof);
  for (int64_t d = 0; d < s.ndof; d++) keep[d] = !exclude[d];
  std::vector<int64_t> ap;
  std::vector<int32_t> ac;
  sym_adjacency(s.ndof, s.rowptr.data(), s.colind.data(), keep, ap, ac);
  finish_tree(s.ndof, ap, ac, s.rowptr.data(), s.colind.data(), nd.parent, dof_node, last_node, T);
}

void lu_analyse_graph(int64_t n, const int32_t *rowptr, const int32_t *colind, const std::vector<uint8_t> &exclude,
                      const std::vector<uint8_t> &last, int leaf_dofs, LuTree &T) {
  if ((int64_t)exclude.size() != n || (int64_t)last.size() != n)
    throw std::invalid_argument("lu: exclude/last must have one entry per row");
  std::vector<uint8_t> keep(n), keep_nd(n);
  for (int64_t i = 0; i < n; i++) {
    keep[i] = !exclude[i];
    keep_nd[i] = !exclude[i] && !last[i];
  }
  std::vector<int64_t> ap, gp;
  std::vector<int32_t> ac, gc;
  sym_adjacency(n, rowptr, colind, keep, ap, ac);
  sym_adjacency(n, rowptr, colind, keep_nd, gp, gc);  // the graph the dissection cuts (no `last` rows)
  std::vector<int32_t> mark(n, -1), lev(n, -1);
  auto bfs = [&](int32_t src, int tag, std::vector<int32_t> &order) {
    order.clear();
    order.push_back(src);
    lev[src] = 0;
    mark[src] = tag;
    for (size_t q = 0; q < order.size(); q++) {
      const int32_t v = order[q];
      for (int64_t k = gp[v]; k < gp[v + 1]; k++) {
        const int32_t w = gc[k];
        if (mark[w] != tag - 1) continue;  // outside the set, or already reached
        mark[w] = tag;
        lev[w] = lev[v] + 1;
        order.push_back(w);
      }
    }
  };
  int tagc = 0, cstamp = 0;
  std::vector<int32_t> cvis(n, -1);
  auto split = [&](const std::vector<int32_t> &V, int, std::vector<int32_t> &A, std::vector<int32_t> &B,
                   std::vector<int32_t> &S) -> bool {
    // mark the set, BFS from V[0] to a far vertex, BFS again from it; cut at the level holding the median
    const int t0 = (tagc += 3);
    for (int32_t v : V) mark[v] = t0 - 1;
    std::vector<int32_t> order;
    bfs(V[0], t0, order);
    if (order.size() < V.size()) {
      // disconnected: whole components to either side (largest first, to the lighter side), no separator needed
      const int cs = 2 * (++cstamp);
      for (int32_t v : V) cvis[v] = cs;  // in the set, not visited
      std::vector<std::vector<int32_t>> comps;
      for (int32_t v0 : V) {
        if (cvis[v0] != cs) continue;
        std::vector<int32_t> comp{v0};
        cvis[v0] = cs + 1;
        for (size_t h = 0; h < comp.size(); h++) {
          const int32_t v = comp[h];
          for (int64_t k = gp[v]; k < gp[v + 1]; k++)
            if (cvis[gc[k]] == cs) {
              cvis[gc[k]] = cs + 1;
              comp.push_back(gc[k]);
            }
        }
        comps.push_back(std::move(comp));
      }
      std::stable_sort(comps.begin(), comps.end(), [](const std::vector<int32_t> &a, const std::vector<int32_t> &b) {
        return a.size() > b.size();
      });
      for (auto &c : comps) {
        std::vector<int32_t> &side = A.size() <= B.size() ? A : B;
        side.insert(side.end(), c.begin(), c.end());
      }
      for (int32_t v : V) mark[v] = -1;
      return true;
    }
    const int32_t far = order.back();
    for (int32_t v : V) mark[v] = t0;  // re-arm: reached vertices carry t0, so t0 + 1 - 1 == t0 marks "in set"
    bfs(far, t0 + 1, order);
    const int maxl = lev[order.back()];
    if (maxl < 2) {
      for (int32_t v : V) mark[v] = -1;
      return false;
    }
    std::vector<int64_t> cnt(maxl + 1, 0);
    for (int32_t v : order) cnt[lev[v]]++;
    const int64_t half = (int64_t)V.size() / 2;
    int cut = 1;
    int64_t acc = cnt[0];
    while (cut < maxl - 1 && acc + cnt[cut] < half) acc += cnt[cut++];
    for (int32_t v : V) {
      if (mark[v] != t0 + 1)
        B.push_back(v);  // not reached: another component, no edge to the reached part
      else if (lev[v] < cut)
        A.push_back(v);
      else if (lev[v] == cut)
        S.push_back(v);
      else
        B.push_back(v);
    }
    for (int32_t v : V) mark[v] = -1;
    return true;
  };
  // dissect only the rows that are neither excluded nor `last`
  std::vector<int32_t> ids;
  for (int64_t i = 0; i < n; i++)
    if (keep_nd[i]) ids.push_back((int32_t)i);
  NdTree nd;
  nd.vnode.assign(n, -1);
  {
    // run the generic dissection on the compacted id list
    struct Task {
      std::vector<int32_t> v;
      int parent;
    };
    std::vector<Task> stack;
    stack.push_back(Task{ids, -1});
    std::vector<int32_t> A, B, S;
    while (!stack.empty()) {
      Task t = std::move(stack.back());
      stack.pop_back();
      if (t.v.empty()) continue;
      A.clear();
      B.clear();
      S.clear();
      if ((int64_t)t.v.size() <= std::max(1, leaf_dofs) || !split(t.v, 0, A, B, S) || (A.empty() && B.empty())) {
        const int node = nd.add(t.parent);
        for (int32_t v : t.v) nd.vnode[v] = node;
        continue;
      }
      const int node = nd.add(t.parent);
      for (int32_t v : S) nd.vnode[v] = node;
      stack.push_back(Task{B, node});
      stack.push_back(Task{A, node});
    }
  }
  int last_node = -1;
  bool any_last = false;
  for (int64_t i = 0; i < n; i++) any_last |= last[i] && !exclude[i];
  if (any_last) {
    last_node = (int)nd.parent.size();
    for (auto &p : nd.parent)
      if (p < 0) p = last_node;
    nd.parent.push_back(-1);
    nd.depth.push_back(-1);
  }
  std::vector<int32_t> dof_node(n, -1);
  for (int64_t i = 0; i < n; i++) {
    if (exclude[i]) continue;
    dof_node[i] = last[i] ? last_node : nd.vnode[i];
  }
  finish_tree(n, ap, ac, rowptr, colind, nd.parent, dof_node, last_node, T);
}

}  // namespace opem
