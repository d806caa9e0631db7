// Symbolic phase of the multifrontal sparse LU: nested-dissection ordering, assembly tree, update sets and the
// index maps the device kernels use.
//
// The reference preconditions its GMRES with a full sparse factorization (PCCHOLESKY, then PCLU on retry:
// src/OpenParEM3D/solveComplexLinearSystem.c:133-168, MUMPS underneath).  This is the ordering and structure
// half of that factorization, written for one GPU: every front is a dense column-major block in one flat device
// buffer, fronts of one tree level are independent, and the maps below let one kernel launch assemble or
// extend-add a whole level.
//
// Ordering.  Nested dissection on a vertex graph: the mesh's vertex graph bisected by coordinate (each cut at the
// median of the longest extent, separator = the smaller one-sided boundary), or, for a caller's matrix without
// geometry, the matrix graph bisected by breadth-first level structures.  A DOF belongs to the deepest tree node
// among its entity's vertices (those vertices are pairwise adjacent, so they lie on one root-to-leaf path and the
// DOF couples only to its node's subtree and ancestors).  Excluded DOFs (Dirichlet rows) are left out; DOFs marked
// `last` (wave-port DOFs) form one root above the rest, so the factorization below that root is shared by every
// excitation and only the root differs per driven port.
#pragma once
#include <cstdint>
#include <vector>

#include "space.h"

namespace opem {

struct LuTree {
  int64_t ndof = 0;                 // rows of the matrix the pattern describes
  int64_t n = 0;                    // factored DOFs
  std::vector<int32_t> perm;        // elimination position -> DOF
  std::vector<int32_t> iperm;       // DOF -> elimination position, -1 when excluded
  int nnodes = 0;                   // tree nodes in postorder (children before parents)
  std::vector<int64_t> dbeg;        // nnodes+1: pivot rows of node t are positions [dbeg[t], dbeg[t+1])
  std::vector<int64_t> uoff;        // nnodes+1: update rows of node t are uidx[uoff[t] .. uoff[t+1])
  std::vector<int32_t> uidx;        // elimination positions, ascending per node
  std::vector<int32_t> umap;        // same layout: row (and column) of each update entry in the parent's front
  std::vector<int32_t> parent;      // -1 for a root
  std::vector<int32_t> level;       // 0 for a leaf, else 1 + the largest child level
  std::vector<int32_t> slot;        // index of the node among its parent's children
  std::vector<int64_t> foff;        // nnodes+1: offset of node t's n_f x n_f front in the flat buffer
  std::vector<int64_t> asm_pos;     // per pattern nonzero: position in the flat front buffer, -1 when excluded
  int nlevels = 0, max_children = 0;
  int last_node = -1;               // the root holding the `last` DOFs, -1 when there are none
  int64_t max_front = 0;
  double flops = 0;                 // real flops of one complex factorization (8 per complex multiply-add)

  int64_t nd(int t) const { return dbeg[t + 1] - dbeg[t]; }
  int64_t nu(int t) const { return uoff[t + 1] - uoff[t]; }
  int64_t nf(int t) const { return nd(t) + nu(t); }
};

// Mesh ordering: nested dissection of the vertex graph of m by coordinate bisection down to leaves of at most
// leaf_vertices vertices.  exclude / last: per DOF of s (size s.ndof).
void lu_analyse_mesh(const Mesh &m, const Space &s, const std::vector<uint8_t> &exclude,
                     const std::vector<uint8_t> &last, int leaf_vertices, LuTree &T);

// Matrix-graph ordering for a caller's CSR pattern (n rows, symmetric structure assumed: the graph is the
// pattern plus its transpose): nested dissection by breadth-first level structures down to leaves of at most
// leaf_dofs DOFs.
void lu_analyse_graph(int64_t n, const int32_t *rowptr, const int32_t *colind, const std::vector<uint8_t> &exclude,
                      const std::vector<uint8_t> &last, int leaf_dofs, LuTree &T);

}  // namespace opem
