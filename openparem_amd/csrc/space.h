// Host-side mesh topology, ND_p DOF numbering, shared CSR pattern and boundary-face matrices.
//
// Replaces what the reference gets from MFEM's ParFiniteElementSpace/ParBilinearForm
// (src/OpenParEM3D/fem3D.cpp:65-123) and the hypre->PETSc merge (csr.c:171-223): one
// sparsity pattern serves K, M, P, Q and every A(omega) of a sweep.
#pragma once
#include <complex>
#include <cstdint>
#include <string>
#include <vector>

namespace opem {

using cplx = std::complex<double>;

struct Mesh {
  int64_t nv = 0, nt = 0, nbf = 0;
  std::vector<double> xyz;         // nv*3
  std::vector<int32_t> tets;       // nt*4 (as given)
  std::vector<int32_t> tet_attr;   // nt
  std::vector<int32_t> bfaces;     // nbf*3
  std::vector<int32_t> bface_attr; // nbf
};

struct Space {
  int p = 0, nd = 0, ndf = 0;
  int k_edge = 0, k_face = 0, k_cell = 0;
  int64_t ndof = 0, nnz = 0;
  std::vector<int32_t> tets_sorted;   // nt*4, vertices ascending
  std::vector<int32_t> bf_sorted;     // nbf*3
  std::vector<int32_t> elem_dofs;     // nt*nd
  std::vector<int32_t> bface_dofs;    // nbf*ndf
  std::vector<int32_t> bface_owner;   // owning tet of each boundary face (first in traversal)
  std::vector<int32_t> rowptr;        // ndof+1
  std::vector<int32_t> colind;        // nnz
  std::vector<int32_t> d2e_ptr;       // ndof+1 : DOF -> element occurrences
  std::vector<int32_t> d2e;           // packed e*nd + local, ascending in e
  std::vector<int32_t> owner_elem;    // per DOF: first element containing it
  std::vector<int32_t> owner_local;   // per DOF: local index in that element
  std::vector<double> gm, gk;         // nt*6 geometric factors (mass, curl-curl)
  std::vector<double> detj;           // nt
};

// Build the space; throws std::runtime_error on bad input.
void build_space(const Mesh &m, int p, Space &s, int nthreads);

// Vertex-patch additive Schwarz structure (Arnold-Falk-Winther H(curl) subspace decomposition):
// patch v = DOFs of every entity (edge, face, cell) incident to vertex v, minus excluded DOFs.
struct Patches {
  int64_t npatch = 0;
  int nmax = 0;
  std::vector<int32_t> idx_off;   // npatch+1, into idx
  std::vector<int32_t> idx;       // global DOFs, ascending per patch
  std::vector<int64_t> val_off;   // npatch+1, into n_p*n_p blocks
  std::vector<int32_t> nz;        // CSR position of (idx[i], idx[j]) or -1, per block
  std::vector<int32_t> d2p_ptr;   // ndof+1
  std::vector<int32_t> d2p_y;     // position in idx of each (DOF, patch-row) pair (its patch-output row)
};
void build_patches(const Space &s, const std::vector<uint8_t> &exclude, Patches &P, int nthreads);

// Position of (row, col) in the CSR pattern, -1 if absent.
int64_t csr_find(const Space &s, int32_t row, int32_t col);

// Unit-coefficient boundary-face matrices (row-major ndf*ndf each) for face f.
void face_matrices(const Mesh &m, const Space &s, int64_t f, double *Kf, double *Mf);

}  // namespace opem
