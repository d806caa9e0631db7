// Device kernels of the per-frequency solve path (gfx950) and their host launchers.
// All launchers are asynchronous on the given stream.
#pragma once
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstdint>

namespace opem {

// Largest Krylov basis (restart + 1) the reduction kernels accept.
constexpr int kMaxBasis = 402;

// ---------------------------------------------------------------- assembly
// Base volume matrices on the shared pattern, one wavefront per row (deterministic gather):
//   K[a][nz], M[a][nz], Q[a][nz] for attribute slot a (unit coefficients).
void launch_assemble_volume(hipStream_t st, int64_t ndof, const int32_t *rowptr, const int32_t *colind,
                            const int32_t *d2e_ptr, const int32_t *d2e, const int32_t *elem_slot,
                            const double *gm, const double *gk, const double *sdet, const double *TM,
                            const double *TK, const double *TQ, int nd, int nslot, int64_t nnz, double *K,
                            double *M, double *Q);

// Element matrices of independent tets (assembly kernel alone), via MFMA contraction of
// per-element geometric factors with the reference tensors.
//   prec 0: FP64 (v_mfma_f64_16x16x4f64), 1: FP32 (v_mfma_f32_16x16x4f32), 2: BF16 inputs
//   (v_mfma_f32_16x16x32_bf16) with per-element scaling, FP32 accumulate.
// geo: ne x 16 factors (6 K, 6 M-real, 4 pad) and ne x 8 (6 M-imag, 2 pad) are formed inside
// from vertices; out: ne x (2*nd*nd) {Re A_e, Im A_e}, FP64 (prec 0) or FP32.
void launch_element_matrices(hipStream_t st, int prec, int p, int64_t ne, const double *verts,
                             const double *er, const double *tand, double k0, const void *tab, void *geo,
                             void *out);
size_t element_tab_bytes(int prec, int p);
void element_table(int prec, int p, void *host_out);
size_t element_geo_bytes(int prec, int64_t ne);

// ---------------------------------------------------------------- per-frequency values
// A_f[nz] = sum_a cK[f][a] K[a][nz] + cM[f][a] M[a][nz]  (complex coefficients, interleaved)
// then A_f[ov_nz[i]] += sum_t cT[f][t] ov_val[t][i].  Writes FP64 and (if a32) FP32 copies.
void launch_combine(hipStream_t st, int nf, int64_t nnz, int nslot, const double *cK, const double *cM,
                    const double *K, const double *M, int64_t nov, int nterm, const int32_t *ov_nz,
                    const double *ov_val, const double *cT, double *a64, float *a32);

// ---------------------------------------------------------------- SpMV
// For matrix group g (values vals + g*vstride; vstride 0 = one shared matrix) and its R right-hand sides (systems g*R..g*R+R-1):
//   y_s = rowmask_s .* (A_g x_s)          (mode 0)
//   y_s = rowmask_s .* (b_s - A_g x_s)    (mode 1, residual)   b, y may alias nothing
// rowmask: per excitation r in [0,R): uint8 N (1 = free row).  Vectors: s-major, N complex.
template <typename V, typename X>
void launch_spmv(hipStream_t st, int ngroups, int R, int64_t n, const int32_t *rowptr, const int32_t *colind,
                 const V *vals, int64_t nnz, const uint8_t *rowmask, const X *x, X *y, const X *b,
                 const int32_t *active, int mode, int lanes_per_row, int64_t vstride);
// One explicit complex CSR matrix and one vector: y = rowmask .* (A x), rowmask null = every row free.
// Accumulates in FP64 for FP32 storage too.  row_group G in {1, 2, 3}: G consecutive rows share one column list
// (csr_row_group must have returned G for this pattern).  lanes_per_row: 4..64 for G = 1, 8/16/32 for G > 1.
template <typename V>
void launch_csr_spmv(hipStream_t st, int64_t n, const int32_t *rowptr, const int32_t *colind, const V *vals,
                     const uint8_t *rowmask, const V *x, V *y, int row_group, int lanes_per_row);
// Largest G in {3, 2} for which every G consecutive rows (from row 0) share one column list, else 1.
int csr_row_group(int64_t n, const int32_t *rowptr, const int32_t *colind);

// ---------------------------------------------------------------- batched sweep SpMV
// The S systems of a batch on one pattern: A_s = Kt + c_s (Mr - j Mi) + sum_t d_{s,t} T_t, with Kt, Mr, Mi
// frequency independent (Mi null when every material is lossless) and the boundary terms T_t stored on
// their rows only (ov_ptr n+1, ov_col, ov_val[t][nov]).
template <typename V>
struct SweepMatrix {
  int64_t n, nnz;
  const int32_t *rowptr, *colind;
  const V *Kt, *Mr, *Mi;
  const int32_t *ov_ptr, *ov_col;
  const V *ov_val;
  int64_t nov;
  int nterm;
  int blk16_max;  // largest number of entries in 16 consecutive rows (LDS staging bound)
};
// y[s*n + row] = rowmask[s%R][row] * (mode ? -(A_s x_s) : A_s x_s)[row] for active s; x system-interleaved
// (x[j*S + s]); csys[s] = c_s, dsys[s][t] complex.  S <= 256.
// form (S > 8 only; S <= 8 always runs the narrow form): 0 = the fixed rule (paired tiles when S is even and
// OPEM_SPMV_PAIR allows, else one-system-per-lane tiles, or the wide form under OPEM_SPMV_WIDE), 1 = paired tiles,
// 2 = one-system-per-lane tiles, 3 = one row per wavefront.  Every form walks a row's entries in the same order.
template <typename V, typename X>
void launch_spmv_sweep(hipStream_t st, const SweepMatrix<V> &M, int S, int R, const V *csys, const V *dsys,
                       const uint8_t *rowmask, const X *x, X *y, int mode, const int32_t *active, int lpr,
                       int form = 0);
// Whether form 1..3 can run this batch (the paired form needs S even and its staged rows within 64 KB of LDS).
template <typename V>
inline bool spmv_sweep_form_ok(const SweepMatrix<V> &M, int S, int form) {
  if (S <= 8 || S > 256) return false;
  if (form == 1)
    return S % 2 == 0 && M.blk16_max > 0 &&
           (size_t)M.blk16_max * ((M.Mi ? 3 : 2) * sizeof(V) + 4) <= 64 * 1024;
  return form == 2 || form == 3;
}
// OPEM_SPMV_TUNE: 1 (default) times the forms of the batched operator once per batch shape and context and keeps
// the fastest whose output is bit-identical to the fixed rule's; 0 keeps the fixed rule; 2 as 1, printing timings.
// When OPEM_SPMV_PAIR or OPEM_SPMV_WIDE is set and OPEM_SPMV_TUNE is not, tuning is off (the forced form runs).
int spmv_tune_level();
// out += sum_i (2 i + 1) a[i] (mod 2^64) over count 64-bit words: a position-weighted checksum.
void launch_hash64(hipStream_t st, int64_t count, const uint64_t *a, unsigned long long *out);
// out64 = [Kt | Mr | Mi] (nnz each; Mi only if lossy), Kt = sum_a w[a][0] K_a, Mr = sum_a w[a][1] M_a,
// Mi = sum_a w[a][2] M_a; out32 (optional) the same in FP32.
void launch_base_matrices(hipStream_t st, int64_t nnz, int nslot, const double *K, const double *M, const double *w,
                          double *out64, float *out32, int lossy);
// system-major [S][n] <-> system-interleaved [n][S]
template <typename X>
void launch_transpose(hipStream_t st, int64_t n, int S, const X *in, X *out, bool to_il);
// x += z, both system-interleaved, for active systems
template <typename X, typename Z>
void launch_axpy_il(hipStream_t st, int S, int64_t n, const Z *z, X *x, const int32_t *active);

// ---------------------------------------------------------------- polynomial preconditioner
// One step of p(B) v for every system (vectors system-major, stride n): u += c_s prod, then (t != null)
// prod -= c_s t, with c_s = coef[s*ldc], coef[s*ldc+1] (complex); first: u = c_s prod.
template <typename X>
void launch_poly_step(hipStream_t st, int S, int64_t n, const double *coef, int ldc, X *prod, const X *t, X *u,
                      bool first, const int32_t *active);
// out[s*ldout + i] = in[s*ldin + i], i < n, for active s
template <typename X>
void launch_copy_sys(hipStream_t st, int S, int64_t n, const X *in, int64_t ldin, X *out, int64_t ldout,
                     const int32_t *active);

// ---------------------------------------------------------------- Krylov helpers (batched, S systems)
// partial[s][blk][i] = sum_{chunk} conj(Vb[s][i][n]) w[s][n], i < k;  vectors in basis stride ldv;
// nblk must be mdot_blocks(n): one partial per wavefront (4 per workgroup of 2048 elements).
template <typename X>
void launch_mdot(hipStream_t st, int S, int64_t n, int k, const X *Vb, int64_t ldv_sys, const X *w,
                 int64_t ldw_sys, double *partial, int nblk, const int32_t *active);
// w[s] -= sum_i Vb[s][i] h[s][i]  ; optionally partial norms^2 of the result into part2[s][blk]
template <typename X>
void launch_mupdate(hipStream_t st, int S, int64_t n, int k, const X *Vb, int64_t ldv_sys, X *w,
                    int64_t ldw_sys, const double *h, int ldh, double *part2, int nblk, const int32_t *active);
// out[s][i] = sum_blk partial[s][blk][i]  (fixed order) ; if acc: Hcol[s][i] += out
void launch_reduce(hipStream_t st, int S, int nblk, int k, const double *partial, double *out, int ldo,
                   double *Hcol, int ldH, const int32_t *active);
// as launch_reduce with ldp complex entries per partial block (launch_reduce: ldp = kMaxBasis)
void launch_reduce_ld(hipStream_t st, int S, int nblk, int k, int ldp, const double *partial, double *out, int ldo,
                      double *Hcol, int ldH, const int32_t *active);

// ---------------------------------------------------------------- s-step (block) Arnoldi, ns = 2..4
// Z holds the ns new vectors, z_c of system s at Z + (c*S + s)*n.
// partial[s][blk][v][c] = conj(V[s][v]) . z_c[s] over a chunk (one pass over V; nblk = mdot_blocks(n));
// reduce with launch_reduce_ld(k = ns*k, ldp = ns*kMaxBasis, ldo = ns*kMaxBasis) into C[s][v][c].
template <typename X>
void launch_mdot_s(hipStream_t st, int ns, int S, int64_t n, int k, const X *Vb, int64_t ldv_sys, const X *Z,
                   double *partial, int nblk, const int32_t *active);
// z_c -= sum_v V[v] C[v][c] (C row stride ldc = kMaxBasis), Gram partials gram[s][blk][ns][ns] (complex, upper
// triangle) of the result (nblk = reduce_blocks(n))
template <typename X>
void launch_mupdate_s(hipStream_t st, int ns, int S, int64_t n, int k, const X *Vb, int64_t ldv_sys, X *Z,
                      const double *C, int ldc, double *gram, int nblk, const int32_t *active);
// out[s][e] = sum_blk gram[s][blk][e], e < ne
void launch_reduce_small(hipStream_t st, int S, int nblk, int ne, const double *gram, double *out,
                         const int32_t *active);
// Cholesky QR: R[s] (ns x ns complex, R^H R = Gram) and q_c = column c of Z R^{-1}, written at Q + s*ldq_sys + c*n
template <typename X>
void launch_cholqr_s(hipStream_t st, int ns, int S, int64_t n, const double *gram, const X *Z, X *Q, int64_t ldq_sys,
                     double *R, const int32_t *active);
// raw Hessenberg columns j .. j+ns-1 into Hraw and Hw; nrm2[t*S + s] = Givens subdiagonal of column j+t, squared
void launch_hcols_s(hipStream_t st, int ns, int S, int m, int j, const double *C, int ldc, const double *R,
                    double *Hraw, double *Hw, double *nrm2, const int32_t *active);

// v[s] = w[s] / nrm[s]  ; nrm from nrm2[s] (sqrt taken inside), also stored into H[s][(k+1)*ldH? ]
template <typename X>
void launch_scale(hipStream_t st, int S, int64_t n, const X *w, int64_t ldw_sys, X *v, int64_t ldv_sys,
                  const double *nrm2, const int32_t *active);

// ---------------------------------------------------------------- block Jacobi
// Build owned-row inverse blocks for systems s (matrix group s/R, mask r = s%R).
// elem_nz[e][a*nd+b] = CSR position of (dof_a, dof_b); a64 = FP64 values of all matrix groups.
template <typename P>
void launch_bj_build(hipStream_t st, int S, int R, int nd, const int32_t *elem_dofs, const int32_t *elem_nz,
                     const double *a64, int64_t nnz, const uint8_t *rowmask, int64_t n, const int32_t *owner_elem,
                     P *inv_rows, const int32_t *elem_of_block, int64_t nblocks);
// z[r*S + s] = sum_b inv_rows[s][r][b] v[s][elem_dofs[owner_elem[r]][b]]  (z system-interleaved)
template <typename P, typename X>
void launch_bj_apply(hipStream_t st, int S, int64_t n, int nd, const int32_t *elem_dofs,
                     const int32_t *owner_elem, const P *inv_rows, const X *v, int64_t ldv_sys, X *z,
                     const int32_t *active, int64_t inv_stride);

// ---------------------------------------------------------------- vertex-patch additive Schwarz
// Device view of the patch structure.  Patches without port DOFs ("shared") have one inverse per
// preconditioner group; patches touching a port DOF have one per excitation (their Dirichlet rows
// differ).  A preconditioner group is pg consecutive frequency groups of the batch; its inverses are
// those of the mean of the groups' matrices.
struct PatchLayout {
  const int32_t *idx_off, *idx;      // patch DOF lists
  const int64_t *val_off;            // offset of the block in its region (FP64 n_p^2, FP16 n_p*npk)
  const int64_t *nz_off;             // offset of the block's CSR-position map in pnz
  const uint8_t *port;               // per patch: excitation-dependent
  const int32_t *pnz;                // CSR position of (idx[i], idx[j]) or -1
  const int32_t *d2p_ptr, *d2p_y;    // DOF -> its rows of the patch-output buffer (positions in idx)
  const int64_t *yoff;               // offset of patch p's rows in the per-system patch-output buffer
  int64_t group_stride, shared_size, exc_size, npatch, ytot;
  int nmax;
  int pg;                            // frequency groups per preconditioner group
  int real_inv;                      // FP16 store of a real matrix's inverse: the Im planes are zero, not read
};
// inv: per-preconditioner-group storage (group_stride elements each); FP16 storage takes per-block
// scales.  G = number of frequency groups in the batch (systems s = g*R + r).
// out[q] = mean of a64[g] over the pg frequencies g of group q (nnz complex values each)
void launch_group_mean(hipStream_t st, int G, int pg, int64_t nnz, const double *a64, double *out);
template <typename P>
void launch_patch_build(hipStream_t st, const PatchLayout &L, int G, int R, const double *a64, int64_t nnz,
                        const uint8_t *rowmask, int64_t n, P *inv, float *scale);
// z[d*S + s] = sum_{patches p containing d} (inv_p v_p)[d] (z system-interleaved, v system-major with stride
// ldv_sys); ybuf holds ytot * S patch outputs of type X (system-interleaved).
template <typename P, typename X>
void launch_patch_apply(hipStream_t st, const PatchLayout &L, int ngroups, int R, int64_t n, const P *inv,
                        const float *scale, const X *v, int64_t ldv_sys, X *z, const int32_t *active, void *ybuf);

// ---------------------------------------------------------------- coarse space (two-level Schwarz)
// rc[s][a] = sum_i C_ia (v[s][i] - t[s][i])   (C^T by coarse rows; t may be null)
template <typename X>
void launch_coarse_restrict(hipStream_t st, int S, int nc, const int32_t *col_ptr, const int32_t *col_row,
                            const double *col_val, const X *v, int64_t ldv_sys, const X *t, int64_t ldt_sys,
                            double *rc, const int32_t *active);
// yc[s] = Ainv[s / R] rc[s]   (dense nc x nc complex per matrix group)
void launch_coarse_solve(hipStream_t st, int S, int R, int nc, const double *Ainv, const double *rc, double *yc,
                         const int32_t *active);
// z[i*S + s] += sum_a C_ia yc[s][a]  (z system-interleaved)
template <typename X>
void launch_coarse_prolong(hipStream_t st, int S, int64_t n, int nc, const int32_t *row_ptr, const int32_t *row_col,
                           const double *row_val, const double *yc, X *z, const int32_t *active);

// ---------------------------------------------------------------- misc vector ops (batched)
template <typename X, typename Y>
void launch_convert(hipStream_t st, int64_t count, const X *in, Y *out);
// x[s] += alpha * z[s] (alpha = 1) for active systems
template <typename X, typename Z>
void launch_axpy(hipStream_t st, int S, int64_t n, const Z *z, X *x, const int32_t *active);
// partial norms^2: part[s][blk]
template <typename X>
void launch_norm2(hipStream_t st, int S, int64_t n, const X *x, double *part, int nblk, const int32_t *active);
int reduce_blocks(int64_t n);
int mdot_blocks(int64_t n);
// OPEM_XCD_REMAP=1 remaps block ids so each XCD works on one contiguous range (measured slower on the
// sweep: off by default).
int xcd_remap_enabled();
// OPEM_PATCH_MFMA=0 applies FP16 patch inverses with vector FMAs instead of MFMA (default 1).
int patch_mfma_enabled();
// OPEM_SPMV_WIDE=1 runs the batched operator one row per wavefront over all systems instead of 16-system tiles.
int spmv_wide_enabled();
// OPEM_SPMV_PAIR=0 disables the two-systems-per-lane form of the batched operator.
int spmv_pair_enabled();

// ---------------------------------------------------------------- Givens / Hessenberg (device)
// Apply previous rotations to column j of H[s] ((m+1) x m, column-major, complex), new rotation,
// update g; est[s] = |g[j+1]|; deactivate systems whose est <= thresh[s] (records kfinal[s] = j+1), and every
// active system at jlast, the cycle's last step (jlast <= m - 1).
// nrm2[s] = |w|^2 of the new Arnoldi vector (becomes H[j+1][j]).
void launch_givens_n(hipStream_t st, int S, int m, int j, int jlast, double *H, double *cs, double *sn, double *g,
                     double *est, const double *thresh, int32_t *active, int32_t *kfinal, const double *nrm2);
// y[s] = R^{-1} g  (k = kfinal[s] or j+1 for still active)
void launch_backsolve(hipStream_t st, int S, int m, const double *H, const double *g, const int32_t *kfinal,
                      double *y, const int32_t *cycle_active);
// u[s] = sum_{i<k_s} V[s][i] y[s][i]
template <typename X>
void launch_combine_basis(hipStream_t st, int S, int64_t n, int m, const X *Vb, int64_t ldv_sys,
                          const double *y, const int32_t *kfinal, X *u, const int32_t *cycle_active);

// ---------------------------------------------------------------- port projections
// out[s][k] = sum_b conj(u[s][off_k + b]) x[s][pd[off_k + b]]  (ports concatenated with offsets)
void launch_port_dots(hipStream_t st, int S, int nport, const int32_t *pd_off, const int32_t *pd,
                      const double *u, int64_t ldu, const double *x, int64_t ldx, double *out);

}  // namespace opem
