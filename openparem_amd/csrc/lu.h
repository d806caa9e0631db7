// Numeric multifrontal sparse LU on the GPU (factorization-strength preconditioner of the Krylov solve).
//
// The reference preconditions KSPGMRES with a full sparse Cholesky, then LU on retry
// (src/OpenParEM3D/solveComplexLinearSystem.c:133-168).  This is the MI355X counterpart: the assembly tree of
// lu_symbolic.h, factored level by level, for several matrices with one pattern at once (the frequencies of a
// sweep).  Fronts are dense column-major blocks in one flat device buffer per matrix.  After a front's partial LU
// (row pivoting inside its pivot block), its leading block holds the inverses of its unit-lower and upper factors,
// so every triangular solve of the preconditioner apply is a matrix-vector product that all threads share.
//
// Wave ports.  The tree's `last` root holds every port DOF.  Its front, after the rest of the tree is factored,
// is the Schur complement of the port rows; excitation e (port e driven, its rows Dirichlet) factors only the rows
// of the other ports.  So the matrices of all excitations of one frequency share everything below the root.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "devbuf.h"
#include "lu_symbolic.h"

namespace opem {

struct LuFactorStats {
  double seconds_factor = 0;
  int64_t zero_pivots = 0;  // pivots that were exactly zero (replaced by a tiny value; the factor is then inexact)
};

class LuFactor {
 public:
  LuFactor();
  ~LuFactor();
  LuFactor(const LuFactor &) = delete;
  LuFactor &operator=(const LuFactor &) = delete;

  // Upload the tree and build the per-level task lists.  root_group: one entry per DOF of the `last` root (its
  // pivot rows in order), the excitation that drives it (0 .. nexc-1); unused when the tree has no `last` root.
  void setup(const LuTree &T, int nexc, const std::vector<int32_t> &root_group, hipStream_t st);

  // Factor ninst matrices (complex FP64 values on the device, [ninst][vals_stride], laid out on the pattern the
  // tree was analysed on).  prec 0: FP64 fronts, 1: FP32 fronts.  symmetric: the caller guarantees A = A^T (the
  // complex-symmetric Maxwell matrices); then FP32 factors (FP64 too under OPEM_LU_SYM=2) compute the lower
  // triangles of the trailing updates only and take pivots inside each 128-column outer block.
  void factor(int prec, int ninst, const double *vals, int64_t vals_stride, LuFactorStats *stats,
              bool symmetric = false);

  // z = A^{-1} v for S = ninst * nrhs systems: system s uses instance s / nrhs and excitation s % nexc.
  // v: system-major, v[s * ldv + dof]; z: system-interleaved, z[dof * S + s] (rows outside the factored set and the
  // driven port's rows are zero).  X = double2 or float2.
  // active (optional, device, ninst * nrhs): per system; an instance none of whose systems is active is skipped
  // (its z entries are left zero), so converged systems cost no factor traffic.
  template <typename X>
  void solve(int nrhs, const X *v, int64_t ldv, X *z, const int32_t *active = nullptr);

  // Accumulated kernel timings (HIP events, only while OPEM_LU_TIMING is set in the environment) and counts:
  // out[0..2] trailing-update seconds, real flops, launches; out[3..4] panel seconds, launches; out[5..7] apply
  // matrix-product seconds, matrix bytes read, launches; out[8] factorization wall seconds; out[9] matrices
  // factored; out[10] matrix-applies; out[11] 1 when event timing is on; out[12..14] seconds, real flops and
  // launches of the deep FP32 trailing updates on the 128-tile kernel (out[0..2] count the 64-tile kernel only).
  // reset clears them afterwards.
  void kernel_stats(double *out, bool reset);

  int prec() const { return prec_; }
  bool symmetric() const { return sym_; }  // the last factorization used the symmetric form
  int ninst() const { return ninst_; }
  const LuTree *tree() const { return T_; }
  // Device bytes one factored matrix needs (fronts, port blocks, inverse scratch, pivots, and the apply's vectors
  // and partial sums for nrhs right-hand sides) at precision prec; valid after setup.
  size_t device_bytes_per_instance(int prec, int nrhs) const;

  struct Impl;  // device-side state (lu.hip)

 private:
  Impl *d_;
  const LuTree *T_ = nullptr;
  int prec_ = 0, ninst_ = 0;
  bool sym_ = false;
};

}  // namespace opem
