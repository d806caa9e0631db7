// Numeric multifrontal sparse LU on the GPU: factorization and preconditioner apply (see lu.h).
//
// Factorization, level by level of the assembly tree (all fronts of a level are independent):
//   1. every front is cleared and receives its original matrix entries (one scatter over the pattern);
//   2. children's update blocks are added into their parents (one launch per child slot, so no two writers meet);
//   3. blocked right-looking LU of every front of the level at once, kNB pivot columns per step: panel
//      factorization with row pivoting inside the pivot block, the panel's interchanges on the other columns,
//      the block row of U, and the tiled trailing update (which also forms the Schur complement of the update
//      rows); then the unit-lower and upper factors of each pivot block are inverted (the same triangular
//      block-row and tiled update kernels on scratch) and written back into the pivot block.
// Each kernel launch serves every front of a level and every matrix of the batch (blockIdx.y), so a level costs a
// few launches per kNB columns of its largest pivot block however many fronts it has.
// Apply (multi right-hand side, all levels): forward through the tree with front vectors (children's update
// parts added into the parent's), backward with the solved ancestors' values gathered through the update lists.
// Every dense step is y = M x with M a block of a factored front, so all threads of a level work at once.
#include "lu.h"

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstdio>
#include <map>
#include <tuple>
#include <stdexcept>
#include <type_traits>

namespace opem {

#define LU_LAUNCH_CHECK()                                                                        \
  do {                                                                                           \
    hipError_t e_ = hipGetLastError();                                                           \
    if (e_ != hipSuccess) throw std::runtime_error(std::string("lu kernel launch: ") + hipGetErrorString(e_)); \
  } while (0)

namespace {

constexpr int kNB = 32;            // pivot columns per step of the blocked LU
constexpr int kMaxRhsThread = 8;   // right-hand sides per thread in the apply kernels
constexpr int kKC = 64;            // input rows staged in LDS per pass of the apply kernels

// ---------------------------------------------------------------- complex helpers
__device__ __forceinline__ double2 d2(double2 v) { return v; }
__device__ __forceinline__ double2 d2(float2 v) { return make_double2((double)v.x, (double)v.y); }
template <typename T>
__device__ __forceinline__ T cvt(double2 v);
template <>
__device__ __forceinline__ double2 cvt<double2>(double2 v) {
  return v;
}
template <>
__device__ __forceinline__ float2 cvt<float2>(double2 v) {
  return make_float2((float)v.x, (float)v.y);
}
template <typename T>
__device__ __forceinline__ T cmul_t(T a, T b) {
  T r;
  r.x = a.x * b.x - a.y * b.y;
  r.y = a.x * b.y + a.y * b.x;
  return r;
}
template <typename T>
__device__ __forceinline__ T csub_t(T a, T b) {
  T r;
  r.x = a.x - b.x;
  r.y = a.y - b.y;
  return r;
}
template <typename T>
__device__ __forceinline__ T cinv_t(T a) {
  using R = decltype(a.x);
  // 1/a with a scaling that keeps |a|^2 in range
  const R s = fmax(fabs(a.x), fabs(a.y));
  T r;
  if (s == (R)0) {
    r.x = r.y = (R)0;
    return r;
  }
  const R ax = a.x / s, ay = a.y / s;
  const R den = (ax * ax + ay * ay) * s;
  r.x = ax / den;
  r.y = -ay / den;
  return r;
}
template <typename T>
__device__ __forceinline__ double cabs1(T a) {
  return fabs((double)a.x) + fabs((double)a.y);
}

// Tree arrays on the device, passed by value.
struct LuDev {
  const int64_t *dbeg, *uoff, *foff, *fvoff;
  const int32_t *uidx, *umap, *parent, *kid_off, *kids;
  int64_t n, ppqs, fstride, fvstride;
  const int32_t *iact;  // apply: per instance, 0 when none of its systems is active (its blocks return at once)
};

// ---------------------------------------------------------------- factorization kernels
template <typename T>
__global__ void k_lu_zero(int64_t count, T *__restrict__ a) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x) {
    T z;
    z.x = 0;
    z.y = 0;
    a[i] = z;
  }
}

template <typename T>
__global__ void k_lu_assemble(int64_t nnz, const int64_t *__restrict__ asm_pos, const double2 *__restrict__ vals,
                              int64_t vstride, T *__restrict__ F, int64_t fstride) {
  const int q = blockIdx.y;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nnz; k += (int64_t)gridDim.x * blockDim.x) {
    const int64_t p = asm_pos[k];
    if (p >= 0) F[q * fstride + p] = cvt<T>(vals[q * vstride + k]);
  }
}

// Task: (child, first column of its update block).  The child's update block (rows and columns nd_c .. nf_c-1
// of its front) is added into the parent's front at the rows and columns umap gives.
// sym: the update blocks hold their lower triangles only (symmetric factorization); umap is increasing, so entry
// (a, b) with a >= b lands in the parent's lower triangle, and only those entries are added.
template <typename T>
__global__ __launch_bounds__(256) void k_lu_extend_add(const int2 *__restrict__ tasks, LuDev D, T *__restrict__ F,
                                                       int cb, int sym) {
  const int2 tk = tasks[blockIdx.x];
  const int q = blockIdx.y;
  const int c = tk.x, p = D.parent[c];
  const int64_t ndc = D.dbeg[c + 1] - D.dbeg[c], nuc = D.uoff[c + 1] - D.uoff[c], nfc = ndc + nuc;
  const int64_t nfp = (D.dbeg[p + 1] - D.dbeg[p]) + (D.uoff[p + 1] - D.uoff[p]);
  const T *Fc = F + q * D.fstride + D.foff[c];
  T *Fp = F + q * D.fstride + D.foff[p];
  const int32_t *um = D.umap + D.uoff[c];
  const int64_t b1 = min((int64_t)tk.y + cb, nuc);
  for (int64_t b = tk.y; b < b1; b++) {
    const int64_t pc = (int64_t)um[b] * nfp;
    const T *src = Fc + ndc + nfc * (ndc + b);
    for (int64_t a = (sym ? b : 0) + threadIdx.x; a < nuc; a += blockDim.x) {
      T v = Fp[um[a] + pc];
      const T s = src[a];
      v.x += s.x;
      v.y += s.y;
      Fp[um[a] + pc] = v;
    }
  }
}

// ---------------------------------------------------------------- blocked dense LU of many fronts at once
// A dense block is described by where it lives: selector (0 = fronts, 1 = port blocks, 2 = L-inverse scratch,
// 3 = U-inverse scratch), offset, leading dimension.  Each launch below serves every front of a tree level (and
// every matrix, blockIdx.y); per step of kNB pivot columns: panel LU, row interchanges of the other columns,
// block row of U, trailing update.
constexpr int kPanelThreads = 1024;
constexpr int kTile = 64;
constexpr int kOuter = 128;  // pivot columns per trailing update

struct DenseFront {
  int sel;          // where the front lives (0 fronts, 1 port blocks)
  int nf, nd;       // order, pivot rows
  int pad;
  int64_t off;      // offset of the front (column-major, ld = nf)
  int64_t pp_off;   // offset of its pivot permutation / interchange record (int32 arrays, per-matrix stride)
  int64_t x_off;    // offset of its nd x nd inverse scratch (selectors 2 and 3)
};

struct DenseBase {
  void *p[4];
  int64_t qs[4];  // per-matrix strides (elements)
  int32_t *pp;    // pivot permutations: fronts [matrix][n], port blocks after them
  int32_t *sw;    // interchange record, same layout
  int64_t ppqs;   // per-matrix stride of pp / sw
};

template <typename T>
__device__ __forceinline__ T *dense_ptr(const DenseBase &B, int sel, int64_t off, int q) {
  return (T *)B.p[sel] + q * B.qs[sel] + off;
}

// Panel: columns k0 .. k0+nb-1 over the pivot rows k0 .. nd-1, row pivoting among them.  Columns are taken in
// groups of kGrp: inside a group rank-1 updates, then the group's block row of the later panel columns and one
// rank-kGrp update of them.  Interchanges are applied inside the panel only and recorded in sw; pp follows them.
// Exactly zero pivots are replaced by a tiny value and counted.  The update rows (nd .. nf-1) of the panel are
// solved afterwards by k_dense_lrows.
constexpr int kGrp = 8;

// sym (symmetric factorization): pivots are searched among the rows of the current outer block only, whose
// rows were completed from the lower triangle by k_dense_mirror; the rows below it hold their lower parts only.
template <typename T>
__global__ __launch_bounds__(kPanelThreads) void k_dense_panel(const int32_t *__restrict__ list,
                                                               const DenseFront *__restrict__ fr, DenseBase B, int k0,
                                                               unsigned long long *__restrict__ zero_pivots, int sym) {
  using R = decltype(T().x);
  const DenseFront f = fr[list[blockIdx.x]];
  const int q = blockIdx.y;
  const int nf = f.nf, nd = f.nd;
  const int nb = min(kNB, nd - k0);
  const int pend = sym ? min(nd, (k0 / kOuter + 1) * kOuter) : nd;  // pivot rows searched: j .. pend-1
  T *F = dense_ptr<T>(B, f.sel, f.off, q);
  int32_t *pp = B.pp + q * B.ppqs + f.pp_off;
  int32_t *sw = B.sw + q * B.ppqs + f.pp_off;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  __shared__ double rv[kPanelThreads / 64];
  __shared__ int ri[kPanelThreads / 64];
  __shared__ int piv_s;
  __shared__ T Lg[kGrp * kGrp];
  if (k0 == 0)
    for (int k = tid; k < nd; k += kPanelThreads) pp[k] = k;
  __syncthreads();
  for (int g0 = k0; g0 < k0 + nb; g0 += kGrp) {
    const int g1 = min(g0 + kGrp, k0 + nb);
    for (int j = g0; j < g1; j++) {
      double best = -1.0;
      int bi = j;
      for (int i = j + tid; i < pend; i += kPanelThreads) {
        const double v = cabs1(F[i + (int64_t)nf * j]);
        if (v > best) {
          best = v;
          bi = i;
        }
      }
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        const double v2 = __shfl_xor(best, o, 64);
        const int i2 = __shfl_xor(bi, o, 64);
        if (v2 > best || (v2 == best && i2 < bi)) {
          best = v2;
          bi = i2;
        }
      }
      if (lane == 0) {
        rv[wid] = best;
        ri[wid] = bi;
      }
      __syncthreads();
      if (tid == 0) {
        double bb = rv[0];
        int p = ri[0];
        for (int w = 1; w < kPanelThreads / 64; w++)
          if (rv[w] > bb || (rv[w] == bb && ri[w] < p)) {
            bb = rv[w];
            p = ri[w];
          }
        sw[j] = p;
        if (p != j) {
          const int32_t a = pp[j];
          pp[j] = pp[p];
          pp[p] = a;
        }
        piv_s = bb > 0.0 ? p : -1 - p;
        if (!(bb > 0.0)) atomicAdd(zero_pivots, 1ull);
      }
      __syncthreads();
      int p = piv_s;
      const bool zero = p < 0;
      if (zero) p = -1 - p;
      if (p != j && tid < nb) {
        const int c = k0 + tid;
        const T a = F[j + (int64_t)nf * c];
        F[j + (int64_t)nf * c] = F[p + (int64_t)nf * c];
        F[p + (int64_t)nf * c] = a;
      }
      __syncthreads();
      if (zero && tid == 0) {
        T tiny;
        tiny.x = std::is_same<R, float>::value ? (R)1e-30 : (R)1e-290;
        tiny.y = 0;
        F[j + (int64_t)nf * j] = tiny;
      }
      __syncthreads();
      const T inv = cinv_t(F[j + (int64_t)nf * j]);
      for (int i = j + 1 + tid; i < nd; i += kPanelThreads) {
        const T l = cmul_t(F[i + (int64_t)nf * j], inv);
        F[i + (int64_t)nf * j] = l;
        for (int c = j + 1; c < g1; c++) F[i + (int64_t)nf * c] = csub_t(F[i + (int64_t)nf * c], cmul_t(l, F[j + (int64_t)nf * c]));
      }
      __syncthreads();
    }
    if (g1 < k0 + nb) {
      // block row of the later panel columns: rows g0 .. g1-1 = Lgg^-1 (unit lower) rows g0 .. g1-1
      const int gn = g1 - g0;
      for (int e = tid; e < gn * gn; e += kPanelThreads) Lg[e] = F[(g0 + e % gn) + (int64_t)nf * (g0 + e / gn)];
      __syncthreads();
      if (tid < k0 + nb - g1) {
        const int c = g1 + tid;
        T x[kGrp];
#pragma unroll
        for (int r = 0; r < kGrp; r++)
          if (r < gn) x[r] = F[g0 + r + (int64_t)nf * c];
#pragma unroll
        for (int r = 1; r < kGrp; r++)
          if (r < gn) {
#pragma unroll
            for (int s2 = 0; s2 < kGrp; s2++)
              if (s2 < r) x[r] = csub_t(x[r], cmul_t(Lg[r + gn * s2], x[s2]));
            F[g0 + r + (int64_t)nf * c] = x[r];
          }
      }
      __syncthreads();
      // rank-gn update of the later panel columns on the pivot rows below the group
      for (int i = g1 + tid; i < nd; i += kPanelThreads) {
        T l[kGrp];
#pragma unroll
        for (int s2 = 0; s2 < kGrp; s2++)
          if (s2 < gn) l[s2] = F[i + (int64_t)nf * (g0 + s2)];
        for (int c = g1; c < k0 + nb; c++) {
          T acc = F[i + (int64_t)nf * c];
#pragma unroll
          for (int s2 = 0; s2 < kGrp; s2++)
            if (s2 < gn) acc = csub_t(acc, cmul_t(l[s2], F[g0 + s2 + (int64_t)nf * c]));
          F[i + (int64_t)nf * c] = acc;
        }
      }
      __syncthreads();
    }
  }
}

// Panel of the symmetric factorization: pivots are searched among rows k0 .. pend-1 only (pend = the end of the
// outer block, at most 128 rows), so those rows are staged in LDS, one row per thread, and the 32-column
// elimination runs there with three barriers per column instead of global-memory passes over every pivot row.
// Rows from pend on are solved afterwards by k_dense_lrows (x U_kk^-1), exactly as the update rows are.
template <typename T>
__global__ __launch_bounds__(kOuter) void k_dense_panel_blk(const int32_t *__restrict__ list,
                                                            const DenseFront *__restrict__ fr, DenseBase B, int k0,
                                                            unsigned long long *__restrict__ zero_pivots) {
  using R = decltype(T().x);
  const DenseFront f = fr[list[blockIdx.x]];
  const int q = blockIdx.y;
  const int nf = f.nf, nd = f.nd;
  const int nb = min(kNB, nd - k0);
  const int pend = min(nd, (k0 / kOuter + 1) * kOuter), nr = pend - k0;
  T *F = dense_ptr<T>(B, f.sel, f.off, q);
  int32_t *pp = B.pp + q * B.ppqs + f.pp_off;
  int32_t *sw = B.sw + q * B.ppqs + f.pp_off;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  __shared__ T P[kOuter][kNB + 1];
  __shared__ double red_v[kOuter / 64];
  __shared__ int red_i[kOuter / 64];
  if (k0 == 0)
    for (int k = t; k < nd; k += kOuter) pp[k] = k;
  for (int c = 0; c < nb; c++)  // column c of the panel: consecutive threads on consecutive rows
    if (t < nr) P[t][c] = F[(k0 + t) + (int64_t)nf * (k0 + c)];
  __syncthreads();
  for (int j = 0; j < nb; j++) {
    double best = (t >= j && t < nr) ? cabs1(P[t][j]) : -1.0;
    int bi = t;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const double v2 = __shfl_xor(best, o, 64);
      const int i2 = __shfl_xor(bi, o, 64);
      if (v2 > best || (v2 == best && i2 < bi)) {
        best = v2;
        bi = i2;
      }
    }
    if (lane == 0) {
      red_v[w] = best;
      red_i[w] = bi;
    }
    __syncthreads();
    double bb = red_v[0];
    int p = red_i[0];
#pragma unroll
    for (int w2 = 1; w2 < kOuter / 64; w2++)
      if (red_v[w2] > bb || (red_v[w2] == bb && red_i[w2] < p)) {
        bb = red_v[w2];
        p = red_i[w2];
      }
    const bool zero = !(bb > 0.0);
    if (p != j && t < nb) {  // rows j and p trade their panel values, one column per thread
      const T a = P[j][t];
      P[j][t] = P[p][t];
      P[p][t] = a;
    }
    if (t == 0) {
      sw[k0 + j] = k0 + p;
      if (p != j) {
        const int32_t s2 = pp[k0 + j];
        pp[k0 + j] = pp[k0 + p];
        pp[k0 + p] = s2;
      }
      if (zero) atomicAdd(zero_pivots, 1ull);
    }
    __syncthreads();
    if (zero && t == 0) {
      P[j][j].x = std::is_same<R, float>::value ? (R)1e-30 : (R)1e-290;
      P[j][j].y = 0;
    }
    __syncthreads();
    if (t > j && t < nr) {
      const T l = cmul_t(P[t][j], cinv_t(P[j][j]));
      P[t][j] = l;
      for (int c = j + 1; c < nb; c++) P[t][c] = csub_t(P[t][c], cmul_t(l, P[j][c]));
    }
    __syncthreads();
  }
  for (int c = 0; c < nb; c++)
    if (t < nr) F[(k0 + t) + (int64_t)nf * (k0 + c)] = P[t][c];
}

// Update rows of a factored panel: rows nd .. nf-1 of columns k0 .. k0+nb-1 times U_kk^-1 (from the right).
// Task = (front, first row after nd); one thread per row.
// sym: the rows from the end of the outer block on (not pivot candidates there) instead of from nd; task rows
// count from that row.
template <typename T>
__global__ __launch_bounds__(256) void k_dense_lrows(const int2 *__restrict__ tasks, const DenseFront *__restrict__ fr,
                                                     DenseBase B, int k0, int sym) {
  const int2 tk = tasks[blockIdx.x];
  const DenseFront f = fr[tk.x];
  const int q = blockIdx.y;
  const int nf = f.nf, nd = sym ? min(f.nd, (k0 / kOuter + 1) * kOuter) : f.nd;
  const int nb = min(kNB, f.nd - k0);
  T *F = dense_ptr<T>(B, f.sel, f.off, q);
  __shared__ T Us[kNB * kNB];
  for (int e = threadIdx.x; e < nb * nb; e += 256) Us[e] = F[(k0 + e % nb) + (int64_t)nf * (k0 + e / nb)];
  __syncthreads();
  const int i = nd + tk.y + threadIdx.x;
  if (i >= nf) return;
  T x[kNB];
#pragma unroll
  for (int c = 0; c < kNB; c++)
    if (c < nb) x[c] = F[i + (int64_t)nf * (k0 + c)];
#pragma unroll
  for (int c = 0; c < kNB; c++)
    if (c < nb) {
#pragma unroll
      for (int s2 = 0; s2 < kNB; s2++)
        if (s2 < c) x[c] = csub_t(x[c], cmul_t(x[s2], Us[s2 + nb * c]));
      x[c] = cmul_t(x[c], cinv_t(Us[c + nb * c]));
      F[i + (int64_t)nf * (k0 + c)] = x[c];
    }
}

// Row interchanges of a panel applied to every other column of the front.  Task = (front, first column); one
// thread per column, the nb interchanges in order.
template <typename T>
__global__ __launch_bounds__(256) void k_dense_laswp(const int2 *__restrict__ tasks, const DenseFront *__restrict__ fr,
                                                     DenseBase B, int k0) {
  const int2 tk = tasks[blockIdx.x];
  const DenseFront f = fr[tk.x];
  const int q = blockIdx.y;
  const int c = tk.y + threadIdx.x;
  const int nb = min(kNB, f.nd - k0);
  if (c >= f.nf || (c >= k0 && c < k0 + nb)) return;
  T *F = dense_ptr<T>(B, f.sel, f.off, q) + (int64_t)f.nf * c;
  const int32_t *sw = B.sw + q * B.ppqs + f.pp_off;
  for (int j = k0; j < k0 + nb; j++) {
    const int p = sw[j];
    if (p != j) {
      const T a = F[j];
      F[j] = F[p];
      F[p] = a;
    }
  }
}

// Triangular block-row solve: X (nb x ncols, rows of one block) = T^-1 X, T the nb x nb diagonal block, unit lower
// (upper == 0) or upper with its diagonal (upper == 1).  Task = (job, first column); one thread per column.
struct TrsmJob {
  int tsel, xsel;
  int nb, ncols;
  int64_t toff, xoff;
  int ldt, ldx;
  int upper, pad;
};

template <typename T>
__global__ __launch_bounds__(256) void k_dense_trsm(const int2 *__restrict__ tasks, const TrsmJob *__restrict__ jobs,
                                                    DenseBase B) {
  const int2 tk = tasks[blockIdx.x];
  const TrsmJob J = jobs[tk.x];
  const int q = blockIdx.y;
  __shared__ T Ts[kNB * kNB];
  const T *Tm = dense_ptr<T>(B, J.tsel, J.toff, q);
  for (int e = threadIdx.x; e < J.nb * J.nb; e += 256) Ts[e] = Tm[(e % J.nb) + (int64_t)J.ldt * (e / J.nb)];
  __syncthreads();
  const int c = tk.y + threadIdx.x;
  if (c >= J.ncols) return;
  T *X = dense_ptr<T>(B, J.xsel, J.xoff, q) + (int64_t)J.ldx * c;
  const int nb = J.nb;
  T x[kNB];
#pragma unroll
  for (int r = 0; r < kNB; r++)
    if (r < nb) x[r] = X[r];
  if (!J.upper) {
#pragma unroll
    for (int r = 1; r < kNB; r++)
      if (r < nb) {
#pragma unroll
        for (int s = 0; s < kNB; s++)
          if (s < r) x[r] = csub_t(x[r], cmul_t(Ts[r + nb * s], x[s]));
      }
  } else {
#pragma unroll
    for (int r = kNB - 1; r >= 0; r--)
      if (r < nb) {
#pragma unroll
        for (int s = 0; s < kNB; s++)
          if (s > r && s < nb) x[r] = csub_t(x[r], cmul_t(Ts[r + nb * s], x[s]));
        x[r] = cmul_t(x[r], cinv_t(Ts[r + nb * r]));
      }
  }
#pragma unroll
  for (int r = 0; r < kNB; r++)
    if (r < nb) X[r] = x[r];
}

// Tiled C -= A B, A m x k, B k x n (any k, taken in slices of kNB), C m x n, for a list of jobs; blocks are numbered job after job
// (prefix of tile counts), each block one kTile x kTile tile of its job, on MFMA (v_mfma_f32_16x16x4f32 or
// v_mfma_f64_16x16x4f64).  Wave w owns columns 16w .. 16w+15 of the tile and all four 16-row blocks.  The MFMA
// computes the transposed product (first operand B^T, second A^T), so D[i][j] = C[row j][col i]: consecutive
// lanes own consecutive rows of the column-major C and its stores are 128-byte runs.  In FP64 a complex product is
// four real ones: Re += Br.Ar - Bi.Ai, Im += Bi.Ar + Br.Ai; in FP32 three: P1 = Br.Ar, P2 = Bi.Ai and
// P3 = (Br + Bi).(Ar + Ai) are accumulated and Re = P1 - P2, Im = P3 - P1 - P2 (the FP32 fronts only feed a
// preconditioner).  Operand layouts of the 16x16x4 MFMA: first operand
// P[i][k] in lane i + 16 k, second Q[k][j] in lane j + 16 k; D[i][j] in register i % 4 of lane j + 16 (i / 4)
// (f32) or register i / 4 of lane j + 16 (i % 4) (f64).
// tri = 1 marks a square trailing update on a front's diagonal (a symmetric Schur complement): launches of the
// symmetric factorization number only its lower tiles (tile row >= tile column).
struct GemmJob {
  int asel, bsel, csel, tri;
  int m, n, k, lda;
  int ldb, ldc, pad2, pad3;
  int64_t aoff, boff, coff;
};

typedef float lu_f4 __attribute__((ext_vector_type(4)));
typedef double lu_d4 __attribute__((ext_vector_type(4)));

// Tile (row, column) of a job's local tile index: row-major over ntc tile columns, or, for a triangular job of a
// triangular launch, the lower tiles row by row (local = r (r + 1) / 2 + c, c <= r).
__device__ __forceinline__ void gemm_tile(int64_t local, int ntc, bool tri, int &tr, int &tc) {
  if (tri) {
    int r = (int)((sqrt(8.0 * (double)local + 1.0) - 1.0) * 0.5);
    while ((int64_t)r * (r + 1) / 2 > local) r--;
    while ((int64_t)(r + 1) * (r + 2) / 2 <= local) r++;
    tr = r;
    tc = (int)(local - (int64_t)r * (r + 1) / 2);
  } else {
    tr = (int)(local / ntc);
    tc = (int)(local % ntc);
  }
}

template <typename T>
struct MfmaAcc;
template <>
struct MfmaAcc<float2> {
  using V = lu_f4;
  static __device__ __forceinline__ V step(float p, float q, V c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(p, q, c, 0, 0, 0);
  }
  static constexpr bool f64 = false;
};
template <>
struct MfmaAcc<double2> {
  using V = lu_d4;
  static __device__ __forceinline__ V step(double p, double q, V c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(p, q, c, 0, 0, 0);
  }
  static constexpr bool f64 = true;
};

template <typename T>
__global__ __launch_bounds__(256) void k_dense_gemm(const GemmJob *__restrict__ jobs, const int64_t *__restrict__ prefix,
                                                    int njobs, DenseBase B, int tri) {
  using M = MfmaAcc<T>;
  using V = typename M::V;
  __shared__ T As[kNB][kTile];
  __shared__ T Bs[kNB][kTile];
  const int64_t b = blockIdx.x;
  const int q = blockIdx.y;
  int lo = 0, hi = njobs - 1;  // last job with prefix <= b
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (prefix[mid] <= b)
      lo = mid;
    else
      hi = mid - 1;
  }
  const GemmJob J = jobs[lo];
  const int64_t local = b - prefix[lo];
  int tr, tc;
  gemm_tile(local, (J.n + kTile - 1) / kTile, tri && J.tri, tr, tc);
  const int r0 = tr * kTile, c0 = tc * kTile;
  const T *A = dense_ptr<T>(B, J.asel, J.aoff, q);
  const T *Bm = dense_ptr<T>(B, J.bsel, J.boff, q);
  T *C = dense_ptr<T>(B, J.csel, J.coff, q);
  const int tid = threadIdx.x;
  const int lane = tid & 63, w = tid >> 6, li = lane & 15, lk = lane >> 4;
  V cr[4], ci[4], c3[4];
#pragma unroll
  for (int rb = 0; rb < 4; rb++) {
    cr[rb] = V{0, 0, 0, 0};
    ci[rb] = V{0, 0, 0, 0};
    c3[rb] = V{0, 0, 0, 0};
  }
  // K in slices of kNB columns staged through registers: every load of the next slice is issued (addresses
  // clamped into the blocks) before the current slice is multiplied, and the padding is masked at the LDS store
  constexpr int kLd = kTile * kNB / 256;
  T ra[kLd], rb_[kLd];
  auto gload = [&](int kc) {
    const int K = min(kNB, J.k - kc);
#pragma unroll
    for (int t = 0; t < kLd; t++) {
      const int e = tid + 256 * t, rr = e % kTile, s = e / kTile;
      ra[t] = A[min(r0 + rr, J.m - 1) + (int64_t)J.lda * (kc + min(s, K - 1))];
    }
#pragma unroll
    for (int t = 0; t < kLd; t++) {
      const int e = tid + 256 * t, s = e % kNB, cc = e / kNB;
      rb_[t] = Bm[kc + min(s, K - 1) + (int64_t)J.ldb * min(c0 + cc, J.n - 1)];
    }
  };
  // FP64 slices are loaded where they are stored (holding the next FP64 slice across the MFMAs spills registers)
  constexpr bool prefetch = !M::f64;
  if (prefetch) gload(0);
  for (int kc = 0; kc < J.k; kc += kNB) {  // C is read and written once
    const int K = min(kNB, J.k - kc);
    if (kc > 0) __syncthreads();
    T zero;
    zero.x = 0;
    zero.y = 0;
    if (prefetch) {
#pragma unroll
      for (int t = 0; t < kLd; t++) {
        const int e = tid + 256 * t, rr = e % kTile, s = e / kTile;
        As[s][rr] = (s < K && r0 + rr < J.m) ? ra[t] : zero;
      }
#pragma unroll
      for (int t = 0; t < kLd; t++) {
        const int e = tid + 256 * t, s = e % kNB, cc = e / kNB;
        Bs[s][cc] = (s < K && c0 + cc < J.n) ? rb_[t] : zero;
      }
    } else {
      for (int e = tid; e < kTile * kNB; e += 256) {
        const int rr = e % kTile, s = e / kTile;
        T v = zero;
        if (s < K && r0 + rr < J.m) v = A[(r0 + rr) + (int64_t)J.lda * (kc + s)];
        As[s][rr] = v;
      }
      for (int e = tid; e < kTile * kNB; e += 256) {
        const int s = e % kNB, cc = e / kNB;
        T v = zero;
        if (s < K && c0 + cc < J.n) v = Bm[kc + s + (int64_t)J.ldb * (c0 + cc)];
        Bs[s][cc] = v;
      }
    }
    __syncthreads();
    if (prefetch && kc + kNB < J.k) gload(kc + kNB);
    const int nk = (K + 3) >> 2;
    for (int kb = 0; kb < nk; kb++) {
      const T bv = Bs[4 * kb + lk][16 * w + li];
      if constexpr (M::f64) {
#pragma unroll
        for (int rb = 0; rb < 4; rb++) {
          const T av = As[4 * kb + lk][16 * rb + li];
          cr[rb] = M::step(bv.x, av.x, cr[rb]);
          cr[rb] = M::step(bv.y, -av.y, cr[rb]);
          ci[rb] = M::step(bv.y, av.x, ci[rb]);
          ci[rb] = M::step(bv.x, av.y, ci[rb]);
        }
      } else {  // three products: cr = Br.Ar, ci = Bi.Ai, c3 = (Br + Bi).(Ar + Ai)
        const float bs = bv.x + bv.y;
#pragma unroll
        for (int rb = 0; rb < 4; rb++) {
          const T av = As[4 * kb + lk][16 * rb + li];
          cr[rb] = M::step(bv.x, av.x, cr[rb]);
          ci[rb] = M::step(bv.y, av.y, ci[rb]);
          c3[rb] = M::step(bs, av.x + av.y, c3[rb]);
        }
      }
    }
  }
  if constexpr (!M::f64) {  // Re = P1 - P2, Im = P3 - P1 - P2
#pragma unroll
    for (int rb = 0; rb < 4; rb++) {
      const V p1 = cr[rb], p2 = ci[rb];
      cr[rb] = p1 - p2;
      ci[rb] = c3[rb] - p1 - p2;
    }
  }
  // the values of C a lane updates are loaded together (clamped addresses; all 16 in FP32, 8 at a time in FP64),
  // then the ones inside the job stored
  constexpr int kHalf = M::f64 ? 2 : 4;
#pragma unroll
  for (int h0 = 0; h0 < 4; h0 += kHalf) {
    T cv[kHalf][4];
#pragma unroll
    for (int rb = 0; rb < kHalf; rb++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int row = min(r0 + 16 * (h0 + rb) + li, J.m - 1);
        const int col = min(c0 + 16 * w + (M::f64 ? 4 * r + lk : 4 * lk + r), J.n - 1);
        cv[rb][r] = C[row + (int64_t)J.ldc * col];
      }
#pragma unroll
    for (int rb = 0; rb < kHalf; rb++) {
      const int row = r0 + 16 * (h0 + rb) + li;
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int col = c0 + 16 * w + (M::f64 ? 4 * r + lk : 4 * lk + r);
        if (row < J.m && col < J.n) {
          T v = cv[rb][r];
          v.x -= cr[h0 + rb][r];
          v.y -= ci[h0 + rb][r];
          C[row + (int64_t)J.ldc * col] = v;
        }
      }
    }
  }
}

// Deep trailing updates of FP32 fronts (C -= A B with k = the outer block): 128 x 128 tiles of C per 256-thread
// block.  K is taken in slices of kBK staged through registers, so the next slice's global loads are in flight
// while the current one is multiplied out of LDS.  Complex products use three real products instead of four:
// with P1 = Br.Ar, P2 = Bi.Ai and P3 = (Br + Bi).(Ar + Ai) accumulated on v_mfma_f32_16x16x4f32, the product is
// Re = P1 - P2, Im = P3 - P1 - P2.  As in k_dense_gemm the MFMA forms the transposed product (first operand B^T,
// second A^T), so each store is a 128-byte run of a column of C.  Wave w owns rows 64 (w & 1) .. +63 and columns
// 64 (w >> 1) .. +63 of the tile: 4 x 4 blocks of 16 x 16 with three accumulators each.  LDS rows are XOR-swizzled
// (columns within and across 16-entry groups) so that the slice stores and the four k-rows one MFMA operand
// load spans fall on different banks.  The flat block index (tiles x matrices) is remapped so that each of the
// 8 XCDs receives a contiguous run of tiles, which share A row panels and B column panels in that XCD's L2.
constexpr int kBT = 128;  // tile edge
constexpr int kBK = 32;   // K slice

__device__ __forceinline__ int swz_a(int s, int r) { return r ^ ((s & 3) << 4); }
__device__ __forceinline__ int swz_b(int s, int c) { return c ^ (s & 15) ^ ((s & 3) << 4); }

__global__ __launch_bounds__(256) void k_dense_gemm_t128(const GemmJob *__restrict__ jobs,
                                                         const int64_t *__restrict__ prefix, int njobs,
                                                         int64_t ntiles, int ninst, DenseBase B, int tri) {
  __shared__ float2 As[kBK][kBT];
  __shared__ float2 Bs[kBK][kBT];
  // XCD-contiguous order of the flat (matrix, tile) index: hardware block b runs on XCD b % 8
  const int64_t total = ntiles * ninst, b = blockIdx.x;
  const int64_t qx = total / 8, rx = total % 8, xcd = b % 8;
  const int64_t flat = xcd * qx + min(xcd, rx) + b / 8;
  const int q = (int)(flat / ntiles);
  const int64_t tile = flat % ntiles;
  int lo = 0, hi = njobs - 1;  // last job with prefix <= tile
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (prefix[mid] <= tile)
      lo = mid;
    else
      hi = mid - 1;
  }
  const GemmJob J = jobs[lo];
  const int64_t local = tile - prefix[lo];
  int tr, tc;
  gemm_tile(local, (J.n + kBT - 1) / kBT, tri && J.tri, tr, tc);
  const int r0 = tr * kBT, c0 = tc * kBT;
  const float2 *A = dense_ptr<float2>(B, J.asel, J.aoff, q);
  const float2 *Bm = dense_ptr<float2>(B, J.bsel, J.boff, q);
  float2 *C = dense_ptr<float2>(B, J.csel, J.coff, q);
  const int tid = threadIdx.x;
  const int lane = tid & 63, w = tid >> 6, li = lane & 15, lk = lane >> 4;
  const int wr = (w & 1) * 64, wc = (w >> 1) * 64;
  lu_f4 p1[4][4], p2[4][4], p3[4][4];  // [column block][row block]
#pragma unroll
  for (int cb = 0; cb < 4; cb++)
#pragma unroll
    for (int rb = 0; rb < 4; rb++) {
      p1[cb][rb] = lu_f4{0, 0, 0, 0};
      p2[cb][rb] = lu_f4{0, 0, 0, 0};
      p3[cb][rb] = lu_f4{0, 0, 0, 0};
    }
  float2 ra[16], rb_[16];
  auto gload = [&](int kc) {
    const int K = min(kBK, J.k - kc);
    // every load is issued (addresses clamped into the blocks) and the padding is masked when the slice is stored
    // to LDS, so the 32 loads of a slice stay in flight during the previous slice's MFMAs
#pragma unroll
    for (int t = 0; t < 16; t++) {  // A: 128 rows x kBK, consecutive threads on consecutive rows
      const int e = tid + 256 * t, rr = e & (kBT - 1), s = e >> 7;
      ra[t] = A[min(r0 + rr, J.m - 1) + (int64_t)J.lda * (kc + min(s, K - 1))];
    }
#pragma unroll
    for (int t = 0; t < 16; t++) {  // B: kBK x 128 columns, consecutive threads on consecutive k
      const int e = tid + 256 * t, s = e & (kBK - 1), cc = e >> 5;
      rb_[t] = Bm[kc + min(s, K - 1) + (int64_t)J.ldb * min(c0 + cc, J.n - 1)];
    }
  };
  gload(0);
  for (int kc = 0; kc < J.k; kc += kBK) {
    const int K = min(kBK, J.k - kc);
    __syncthreads();  // the previous slice has been read
#pragma unroll
    for (int t = 0; t < 16; t++) {
      const int e = tid + 256 * t, rr = e & (kBT - 1), s = e >> 7;
      const bool ok = s < K && r0 + rr < J.m;
      As[s][swz_a(s, rr)] = ok ? ra[t] : make_float2(0.f, 0.f);
    }
#pragma unroll
    for (int t = 0; t < 16; t++) {
      const int e = tid + 256 * t, s = e & (kBK - 1), cc = e >> 5;
      const bool ok = s < K && c0 + cc < J.n;
      Bs[s][swz_b(s, cc)] = ok ? rb_[t] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    if (kc + kBK < J.k) gload(kc + kBK);
#pragma unroll 2
    for (int kb = 0; kb < kBK / 4; kb++) {
      const int s = 4 * kb + lk;
      float2 av[4], bv[4];
      float as[4], bs[4];
#pragma unroll
      for (int rb = 0; rb < 4; rb++) {
        av[rb] = As[s][swz_a(s, wr + 16 * rb + li)];
        as[rb] = av[rb].x + av[rb].y;
      }
#pragma unroll
      for (int cb = 0; cb < 4; cb++) {
        bv[cb] = Bs[s][swz_b(s, wc + 16 * cb + li)];
        bs[cb] = bv[cb].x + bv[cb].y;
      }
#pragma unroll
      for (int cb = 0; cb < 4; cb++)
#pragma unroll
        for (int rb = 0; rb < 4; rb++) {
          p1[cb][rb] = __builtin_amdgcn_mfma_f32_16x16x4f32(bv[cb].x, av[rb].x, p1[cb][rb], 0, 0, 0);
          p2[cb][rb] = __builtin_amdgcn_mfma_f32_16x16x4f32(bv[cb].y, av[rb].y, p2[cb][rb], 0, 0, 0);
          p3[cb][rb] = __builtin_amdgcn_mfma_f32_16x16x4f32(bs[cb], as[rb], p3[cb][rb], 0, 0, 0);
        }
    }
  }
  // D[i][j] of block (cb, rb) = C[row r0 + wr + 16 rb + j][col c0 + wc + 16 cb + i]; lane l, register r: i = 4 (l / 16)
  // + r, j = l % 16.  All 64 values of C a lane updates are loaded at once (clamped addresses), then the ones inside
  // the job are stored.
  float2 cv[4][4][4];
#pragma unroll
  for (int rb = 0; rb < 4; rb++)
#pragma unroll
    for (int cb = 0; cb < 4; cb++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int row = min(r0 + wr + 16 * rb + li, J.m - 1), col = min(c0 + wc + 16 * cb + 4 * lk + r, J.n - 1);
        cv[rb][cb][r] = C[row + (int64_t)J.ldc * col];
      }
#pragma unroll
  for (int rb = 0; rb < 4; rb++) {
    const int row = r0 + wr + 16 * rb + li;
#pragma unroll
    for (int cb = 0; cb < 4; cb++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int col = c0 + wc + 16 * cb + 4 * lk + r;
        if (row < J.m && col < J.n) {
          float2 v = cv[rb][cb][r];
          const float re = p1[cb][rb][r] - p2[cb][rb][r];
          const float im = p3[cb][rb][r] - p1[cb][rb][r] - p2[cb][rb][r];
          v.x -= re;
          v.y -= im;
          C[row + (int64_t)J.ldc * col] = v;
        }
      }
  }
}

// Deep trailing updates of FP32 fronts with split-BF16 products: the same 128 x 128 tiles, staging, XCD order and
// store pattern as k_dense_gemm_t128, but each real product runs as three BF16 products on
// v_mfma_f32_16x16x32_bf16 (x = hi + lo with hi = bf16(x), lo = bf16(x - hi); x y ~ hi_x hi_y + hi_x lo_y +
// lo_x hi_y, the dropped lo_x lo_y being 2^-16 of the product), accumulated in FP32.  The complex product takes
// four real products into two accumulators (Re += Br Ar - Bi Ai, Im += Br Ai + Bi Ar; the minus sign flips the
// BF16 sign bits of Ai): three accumulators per block, as in the three-product form, would not fit the register
// file beside the split operands.  The operands keep
// about 16 significant bits, against 24 in FP32; the factor is a preconditioner inside FP64 refinement, and the
// config-4 sweep needs the same GMRES iterations with operands rounded this way (16.27 vs 16.26 per system).
// Operand layouts of the 16x16x32 MFMA: first operand P[i][k] in lane i + 16 g, element e (k = 8 g + e), second
// Q[k][j] in lane j + 16 g, element e; D as for 16x16x4.  Each K slice of 32 is one MFMA step: lane (li, lk) reads
// 8 consecutive k of its row (A) or column (B) from LDS and splits them in registers.
typedef __bf16 lu_bf8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ void split_bf16(const float (&x)[8], lu_bf8 &hi, lu_bf8 &lo) {
#pragma unroll
  for (int e = 0; e < 8; e++) {
    const __bf16 h = (__bf16)x[e];
    hi[e] = h;
    lo[e] = (__bf16)(x[e] - (float)h);
  }
}

__global__ __launch_bounds__(256) void k_dense_gemm_t128_bf16x3(const GemmJob *__restrict__ jobs,
                                                                const int64_t *__restrict__ prefix, int njobs,
                                                                int64_t ntiles, int ninst, DenseBase B, int tri) {
  __shared__ float2 As[kBK][kBT];
  __shared__ float2 Bs[kBK][kBT];
  const int64_t total = ntiles * ninst, b = blockIdx.x;
  const int64_t qx = total / 8, rx = total % 8, xcd = b % 8;
  const int64_t flat = xcd * qx + min(xcd, rx) + b / 8;
  const int q = (int)(flat / ntiles);
  const int64_t tile = flat % ntiles;
  int lo = 0, hi = njobs - 1;  // last job with prefix <= tile
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (prefix[mid] <= tile)
      lo = mid;
    else
      hi = mid - 1;
  }
  const GemmJob J = jobs[lo];
  const int64_t local = tile - prefix[lo];
  int tr, tc;
  gemm_tile(local, (J.n + kBT - 1) / kBT, tri && J.tri, tr, tc);
  const int r0 = tr * kBT, c0 = tc * kBT;
  const float2 *A = dense_ptr<float2>(B, J.asel, J.aoff, q);
  const float2 *Bm = dense_ptr<float2>(B, J.bsel, J.boff, q);
  float2 *C = dense_ptr<float2>(B, J.csel, J.coff, q);
  const int tid = threadIdx.x;
  const int lane = tid & 63, w = tid >> 6, li = lane & 15, lk = lane >> 4;
  const int wr = (w & 1) * 64, wc = (w >> 1) * 64;
  lu_f4 pr[4][4], pi[4][4];  // [column block][row block]: Re and Im of C's update
#pragma unroll
  for (int cb = 0; cb < 4; cb++)
#pragma unroll
    for (int rb = 0; rb < 4; rb++) {
      pr[cb][rb] = lu_f4{0, 0, 0, 0};
      pi[cb][rb] = lu_f4{0, 0, 0, 0};
    }
  float2 ra[16], rb_[16];
  auto gload = [&](int kc) {
    const int K = min(kBK, J.k - kc);
#pragma unroll
    for (int t = 0; t < 16; t++) {  // A: 128 rows x kBK, consecutive threads on consecutive rows
      const int e = tid + 256 * t, rr = e & (kBT - 1), s = e >> 7;
      ra[t] = A[min(r0 + rr, J.m - 1) + (int64_t)J.lda * (kc + min(s, K - 1))];
    }
#pragma unroll
    for (int t = 0; t < 16; t++) {  // B: kBK x 128 columns, consecutive threads on consecutive k
      const int e = tid + 256 * t, s = e & (kBK - 1), cc = e >> 5;
      rb_[t] = Bm[kc + min(s, K - 1) + (int64_t)J.ldb * min(c0 + cc, J.n - 1)];
    }
  };
  gload(0);
  for (int kc = 0; kc < J.k; kc += kBK) {
    const int K = min(kBK, J.k - kc);
    __syncthreads();  // the previous slice has been read
#pragma unroll
    for (int t = 0; t < 16; t++) {
      const int e = tid + 256 * t, rr = e & (kBT - 1), s = e >> 7;
      const bool ok = s < K && r0 + rr < J.m;
      As[s][swz_a(s, rr)] = ok ? ra[t] : make_float2(0.f, 0.f);
    }
#pragma unroll
    for (int t = 0; t < 16; t++) {
      const int e = tid + 256 * t, s = e & (kBK - 1), cc = e >> 5;
      const bool ok = s < K && c0 + cc < J.n;
      Bs[s][swz_b(s, cc)] = ok ? rb_[t] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    if (kc + kBK < J.k) gload(kc + kBK);
    // the first operands (B^T) of the four column blocks, split once for the four row blocks
    lu_bf8 bh[4][2], bl[4][2];
#pragma unroll
    for (int cb = 0; cb < 4; cb++) {
      float xr[8], xi[8];
#pragma unroll
      for (int e = 0; e < 8; e++) {
        const int s = 8 * lk + e;
        const float2 v = Bs[s][swz_b(s, wc + 16 * cb + li)];
        xr[e] = v.x;
        xi[e] = v.y;
      }
      split_bf16(xr, bh[cb][0], bl[cb][0]);
      split_bf16(xi, bh[cb][1], bl[cb][1]);
    }
#pragma unroll
    for (int rb = 0; rb < 4; rb++) {
      float xr[8], xi[8];
#pragma unroll
      for (int e = 0; e < 8; e++) {
        const int s = 8 * lk + e;
        const float2 v = As[s][swz_a(s, wr + 16 * rb + li)];
        xr[e] = v.x;
        xi[e] = v.y;
      }
      lu_bf8 ahr, alr, ahi, ali;
      split_bf16(xr, ahr, alr);
      split_bf16(xi, ahi, ali);
      const lu_bf8 nhi = -ahi, nli = -ali;
#pragma unroll
      for (int cb = 0; cb < 4; cb++) {
        lu_f4 r = pr[cb][rb], m = pi[cb][rb];
        r = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[cb][0], ahr, r, 0, 0, 0);
        r = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[cb][0], alr, r, 0, 0, 0);
        r = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bl[cb][0], ahr, r, 0, 0, 0);
        r = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[cb][1], nhi, r, 0, 0, 0);
        r = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[cb][1], nli, r, 0, 0, 0);
        r = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bl[cb][1], nhi, r, 0, 0, 0);
        m = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[cb][0], ahi, m, 0, 0, 0);
        m = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[cb][0], ali, m, 0, 0, 0);
        m = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bl[cb][0], ahi, m, 0, 0, 0);
        m = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[cb][1], ahr, m, 0, 0, 0);
        m = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[cb][1], alr, m, 0, 0, 0);
        m = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bl[cb][1], ahr, m, 0, 0, 0);
        pr[cb][rb] = r;
        pi[cb][rb] = m;
      }
    }
  }
  // D[i][j] of block (cb, rb) = C[row r0 + wr + 16 rb + j][col c0 + wc + 16 cb + i]; lane l, register r: i = 4 (l / 16)
  // + r, j = l % 16.  All 64 values of C a lane updates are loaded at once (clamped addresses), then the ones inside
  // the job are stored.
  float2 cv[4][4][4];
#pragma unroll
  for (int rb = 0; rb < 4; rb++)
#pragma unroll
    for (int cb = 0; cb < 4; cb++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int row = min(r0 + wr + 16 * rb + li, J.m - 1), col = min(c0 + wc + 16 * cb + 4 * lk + r, J.n - 1);
        cv[rb][cb][r] = C[row + (int64_t)J.ldc * col];
      }
#pragma unroll
  for (int rb = 0; rb < 4; rb++) {
    const int row = r0 + wr + 16 * rb + li;
#pragma unroll
    for (int cb = 0; cb < 4; cb++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int col = c0 + wc + 16 * cb + 4 * lk + r;
        if (row < J.m && col < J.n) {
          float2 v = cv[rb][cb][r];
          v.x -= pr[cb][rb][r];
          v.y -= pi[cb][rb][r];
          C[row + (int64_t)J.ldc * col] = v;
        }
      }
  }
}

// Scratch X = I and Y = I (nd x nd) for every listed front.  Task = (front, first column of a 16-column chunk).
template <typename T>
__global__ void k_dense_eye(const int2 *__restrict__ tasks, const DenseFront *__restrict__ fr, DenseBase B) {
  const int2 tk = tasks[blockIdx.x];
  const DenseFront f = fr[tk.x];
  const int q = blockIdx.y;
  T *X = dense_ptr<T>(B, 2, f.x_off, q), *Y = dense_ptr<T>(B, 3, f.x_off, q);
  const int c1 = min(tk.y + 16, f.nd);
  for (int c = tk.y; c < c1; c++)
    for (int i = threadIdx.x; i < f.nd; i += blockDim.x) {
      T v;
      v.x = i == c ? 1 : 0;
      v.y = 0;
      X[i + (int64_t)f.nd * c] = v;
      Y[i + (int64_t)f.nd * c] = v;
    }
}

// Pivot block <- strictly lower part of L^-1 (scratch 2) + upper part of U^-1 (scratch 3), same tasks.
template <typename T>
__global__ void k_dense_put_inverses(const int2 *__restrict__ tasks, const DenseFront *__restrict__ fr, DenseBase B) {
  const int2 tk = tasks[blockIdx.x];
  const DenseFront f = fr[tk.x];
  const int q = blockIdx.y;
  const T *X = dense_ptr<T>(B, 2, f.x_off, q), *Y = dense_ptr<T>(B, 3, f.x_off, q);
  T *F = dense_ptr<T>(B, f.sel, f.off, q);
  const int c1 = min(tk.y + 16, f.nd);
  for (int c = tk.y; c < c1; c++)
    for (int i = threadIdx.x; i < f.nd; i += blockDim.x)
      F[i + (int64_t)f.nf * c] = i > c ? X[i + (int64_t)f.nd * c] : Y[i + (int64_t)f.nd * c];
}

// Excitation e's port block: E[i + no*k] = S[O[i] + nfr * O[k]], S the root front (the ports' Schur complement).
// Symmetric factorization: before an outer block [ob0, ob1) is factored, its rows right of the diagonal are
// filled from its columns below it, F[i][j] = F[j][i] for ob0 <= i < ob1, j > i.  The trailing updates of a
// symmetric front compute lower tiles only (the Schur complement of a pivot block does not depend on the row
// interchanges inside it, so it stays symmetric), and the row interchanges stay inside the outer block, so these
// rows are the only upper entries the factorization reads.  Task = (front, first row i0 (a multiple of 64 from
// ob0), first column j0 >= i0); a 64 x 64 tile is transposed through LDS.  Reads are at j > i, writes at i < j:
// no tile writes what another reads.
template <typename T>
__global__ __launch_bounds__(256) void k_dense_mirror(const int4 *__restrict__ tasks,
                                                      const DenseFront *__restrict__ fr, DenseBase B, int ob0) {
  const int4 tk = tasks[blockIdx.x];
  const DenseFront f = fr[tk.x];
  const int q = blockIdx.y;
  const int nf = f.nf, ob1 = min(ob0 + kOuter, f.nd), i0 = tk.y, j0 = tk.z;
  T *F = dense_ptr<T>(B, f.sel, f.off, q);
  __shared__ T tile[64][65];
  const int tid = threadIdx.x;
#pragma unroll 4
  for (int t = 0; t < 16; t++) {  // read F[j][i]: consecutive threads on consecutive j (rows of column i)
    const int e = tid + 256 * t, jj = e & 63, ii = e >> 6;
    const int i = i0 + ii, j = j0 + jj;
    if (i < ob1 && j < nf && j > i) tile[ii][jj] = F[j + (int64_t)nf * i];
  }
  __syncthreads();
#pragma unroll 4
  for (int t = 0; t < 16; t++) {  // write F[i][j]: consecutive threads on consecutive i (rows of column j)
    const int e = tid + 256 * t, ii = e & 63, jj = e >> 6;
    const int i = i0 + ii, j = j0 + jj;
    if (i < ob1 && j < nf && j > i) F[i + (int64_t)nf * j] = tile[ii][jj];
  }
}

// sym: the root front holds its lower triangle only; entry (i, k) is read at (max, min).
template <typename T>
__global__ void k_lu_root_gather(const T *__restrict__ S, int64_t sstride, int nfr, const int32_t *__restrict__ O,
                                 int no, T *__restrict__ E, int64_t estride, int sym) {
  const int q = blockIdx.y;
  const int64_t tot = (int64_t)no * no;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(e % no), k = (int)(e / no);
    const int r = sym ? max(O[i], O[k]) : O[i], c = sym ? min(O[i], O[k]) : O[k];
    E[q * estride + e] = S[q * sstride + r + (int64_t)nfr * c];
  }
}

// ---------------------------------------------------------------- apply kernels
// Work vectors are double2 [matrix][row][nrhs].

template <typename X>
__global__ void k_lu_gather_rhs(int64_t n, int nrhs, const int32_t *__restrict__ perm, const X *__restrict__ v,
                                int64_t ldv, double2 *__restrict__ b, const int32_t *__restrict__ iact) {
  const int q = blockIdx.y;
  if (iact && !iact[q]) return;
  const int64_t tot = n * nrhs;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t pos = e / nrhs;
    const int j = (int)(e % nrhs);
    b[q * tot + e] = d2(v[((int64_t)q * nrhs + j) * ldv + perm[pos]]);
  }
}

template <typename X>
// z at the factored rows: x of an active instance, zero for an instance whose systems are all inactive (the rows
// outside the factored set are cleared by k_lu_zero_list, so every entry of z is written once per apply).
__global__ void k_lu_scatter(int64_t n, int nrhs, int64_t S, const int32_t *__restrict__ perm,
                             const double2 *__restrict__ x, X *__restrict__ z, const int32_t *__restrict__ iact) {
  const int q = blockIdx.y;
  const bool active = !iact || iact[q];
  const int64_t tot = n * nrhs;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t pos = e / nrhs;
    const int j = (int)(e % nrhs);
    z[(int64_t)perm[pos] * S + (int64_t)q * nrhs + j] =
        active ? cvt<X>(x[q * tot + e]) : cvt<X>(make_double2(0.0, 0.0));
  }
}

// z = 0 at the rows outside the factored set (rows[0 .. count)), every system.
template <typename X>
__global__ void k_lu_zero_list(int64_t count, const int32_t *__restrict__ rows, int64_t S, X *__restrict__ z) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < count * S;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = e / S;
    z[(int64_t)rows[r] * S + (e - r * S)] = cvt<X>(make_double2(0.0, 0.0));
  }
}

// Instance q is active when any of its right-hand sides (systems q nrhs .. q nrhs + nrhs - 1) is.
__global__ void k_lu_inst_active(int ninst, int nrhs, const int32_t *__restrict__ act, int32_t *__restrict__ iact) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= ninst) return;
  int a = 0;
  for (int j = 0; j < nrhs; j++) a |= act[(int64_t)q * nrhs + j] != 0;
  iact[q] = a;
}

// Front vectors of the nodes of one level (forward solve): k_lu_fwd_init fills kFvRows rows of a front's vector per
// task (the right-hand side on the pivot rows, zero on the update rows), then k_lu_fwd_add, once per child slot,
// adds kFvRows update rows of one child per task at the rows umap gives (children of one slot have different parents and umap
// is injective, so no two threads add to one entry; the slots are added in a fixed order, so the sums do not
// depend on scheduling).  Task = (node, first row).
constexpr int kFvRows = 32;

__global__ __launch_bounds__(256) void k_lu_fwd_init(const int2 *__restrict__ tasks, LuDev D, int nrhs,
                                                     const double2 *__restrict__ ball, double2 *__restrict__ fvall) {
  const int q = blockIdx.y;
  if (D.iact && !D.iact[q]) return;
  const int2 tk = tasks[blockIdx.x];
  const int t = tk.x;
  const int64_t nd = D.dbeg[t + 1] - D.dbeg[t], nf = nd + (D.uoff[t + 1] - D.uoff[t]);
  double2 *fv = fvall + q * D.fvstride + D.fvoff[t] * nrhs;
  const double2 *b = ball + q * D.n * nrhs + D.dbeg[t] * nrhs;
  const int64_t e1 = min((int64_t)tk.y + kFvRows, nf) * nrhs;
  for (int64_t e = (int64_t)tk.y * nrhs + threadIdx.x; e < e1; e += blockDim.x)
    fv[e] = e < nd * nrhs ? b[e] : make_double2(0.0, 0.0);
}

__global__ __launch_bounds__(256) void k_lu_fwd_add(const int2 *__restrict__ tasks, LuDev D, int nrhs,
                                                    double2 *__restrict__ fvall) {
  const int q = blockIdx.y;
  if (D.iact && !D.iact[q]) return;
  const int2 tk = tasks[blockIdx.x];
  const int c = tk.x, p = D.parent[c];
  const int64_t ndc = D.dbeg[c + 1] - D.dbeg[c], nuc = D.uoff[c + 1] - D.uoff[c];
  const double2 *fc = fvall + q * D.fvstride + (D.fvoff[c] + ndc) * nrhs;
  double2 *fp = fvall + q * D.fvstride + D.fvoff[p] * nrhs;
  const int32_t *um = D.umap + D.uoff[c];
  const int64_t e1 = min((int64_t)tk.y + kFvRows, nuc) * nrhs;
  for (int64_t e = (int64_t)tk.y * nrhs + threadIdx.x; e < e1; e += blockDim.x) {
    const int64_t a = e / nrhs;
    const int j = (int)(e - a * nrhs);
    double2 &dst = fp[(int64_t)um[a] * nrhs + j];
    dst.x += fc[e].x;
    dst.y += fc[e].y;
  }
}

enum { K_FWD_L = 0, K_FWD_L21 = 1, K_BWD_U12 = 2, K_BWD_U = 3, K_ROOT_L = 4, K_ROOT_U = 5 };

struct RootDev {
  int node;              // tree node of the `last` root
  int nexc;
  const int32_t *O;      // [nexc][no_max] free rows of the root front per excitation
  const int32_t *no;     // [nexc]
  int no_max;
  int64_t estride;       // elements per (matrix, excitation) block: no_max^2
  const int32_t *rpp;    // pivot permutations of the excitation blocks: pp + n, [matrix stride ppqs][nexc][no_max]
};

// Apply as a split-K product, deterministic: phase 1 (task = node or excitation, first row, first column, slot)
// writes the partial sums of its rows over one chunk of kKCH columns to its slot; phase 2 (task = node or
// excitation, first row, chunk count, first slot) adds a row tile's slots (four interleaved running sums over the
// chunks, in a fixed order) and applies the output rule.  In phase 1 each thread owns one row and up to
// kMaxRhsThread right-hand sides (rhs j = j0 + rstep * m); matrix loads are issued kUnroll at a time with clamped
// indices, so none is conditional; the inputs are staged kKC rows at a time in LDS.
constexpr int kKCH = 128;  // default columns per phase-1 chunk (OPEM_LU_KCH overrides: 32, 64, 256 or 512)
constexpr int kUnroll = 16;

template <int KIND>
__device__ __forceinline__ void apply_geom(const int4 tk, const LuDev &D, const RootDev &Rt, int64_t &rows,
                                           int64_t &cols, int64_t &ld, int &t, int &e, int64_t &dbeg, int64_t &nd,
                                           int64_t &nf) {
  if (KIND == K_ROOT_L || KIND == K_ROOT_U) {
    e = tk.x;
    t = Rt.node;
    dbeg = D.dbeg[t];
    nd = D.dbeg[t + 1] - dbeg;
    nf = nd;
    rows = cols = Rt.no[e];
    ld = rows;
  } else {
    t = tk.x;
    e = 0;
    dbeg = D.dbeg[t];
    nd = D.dbeg[t + 1] - dbeg;
    nf = nd + (D.uoff[t + 1] - D.uoff[t]);
    ld = nf;
    if (KIND == K_FWD_L21) {
      rows = nf - nd;
      cols = nd;
    } else if (KIND == K_BWD_U12) {
      rows = nd;
      cols = nf - nd;
    } else {
      rows = cols = nd;
    }
  }
}

template <typename T, int KIND, int MAXR>
__global__ __launch_bounds__(256) void k_lu_apply_part(const int4 *__restrict__ tasks, LuDev D, RootDev Rt, int nrhs,
                                                       int ty_n, const T *__restrict__ Fall, const T *__restrict__ Eall,
                                                       const int32_t *__restrict__ ppall, const double2 *__restrict__ fvall,
                                                       const double2 *__restrict__ yall,
                                                       const double2 *__restrict__ tall,
                                                       const double2 *__restrict__ xall,
                                                       const double2 *__restrict__ yrall, double2 *__restrict__ part,
                                                       int64_t part_qstride, int kch) {
  extern __shared__ double2 sin[];  // kKC x nsub staged inputs
  const int q = blockIdx.y;
  if (D.iact && !D.iact[q]) return;
  const int4 tk = tasks[blockIdx.x];
  const int rows_wg = 256 / ty_n;
  const int tx = threadIdx.x % rows_wg, ty = threadIdx.x / rows_wg;
  int t, e;
  int64_t rows, cols, ld, dbeg, nd, nf;
  apply_geom<KIND>(tk, D, Rt, rows, cols, ld, t, e, dbeg, nd, nf);
  int j0 = 0, rstep = 1;
  if (KIND == K_ROOT_L || KIND == K_ROOT_U) {
    j0 = e;
    rstep = Rt.nexc;
  }
  const int nsub = (nrhs - j0 + rstep - 1) / rstep;
  const int my_n = nsub > ty ? (nsub - ty + ty_n - 1) / ty_n : 0;
  const T *M;
  if (KIND == K_ROOT_L || KIND == K_ROOT_U) {
    M = Eall + ((int64_t)q * Rt.nexc + e) * Rt.estride;
  } else {
    const T *F = Fall + q * D.fstride + D.foff[t];
    M = KIND == K_FWD_L21 ? F + nd : KIND == K_BWD_U12 ? F + nf * nd : F;
  }
  const int64_t nrows_all = D.n * nrhs;
  const double2 *fv = fvall + q * D.fvstride;
  const double2 *y = yall + q * nrows_all, *tt = tall + q * nrows_all, *x = xall + q * nrows_all;
  const int32_t *pp = ppall + q * D.ppqs + dbeg;
  const int32_t *rpp = (KIND == K_ROOT_L) ? Rt.rpp + q * D.ppqs + (int64_t)e * Rt.no_max : nullptr;
  const int32_t *O = (KIND == K_ROOT_L || KIND == K_ROOT_U) ? Rt.O + (int64_t)e * Rt.no_max : nullptr;
  const double2 *yr = yrall + (int64_t)q * Rt.no_max * nrhs;
  const int64_t i = tk.y + tx;
  const int64_t ic = i < rows ? i : rows - 1;  // clamped row: loads stay inside the block
  const int64_t c_beg = tk.z, c_end = min((int64_t)tk.z + kch, cols);
  double2 acc[MAXR];
#pragma unroll
  for (int m = 0; m < MAXR; m++) acc[m] = make_double2(0.0, 0.0);
  for (int64_t kc = c_beg; kc < c_end; kc += kKC) {
    const int kn = (int)min((int64_t)kKC, c_end - kc);
    for (int s = threadIdx.x; s < kn * nsub; s += 256) {
      const int kk = s / nsub, m = s % nsub;
      const int j = j0 + m * rstep;
      const int64_t k = kc + kk;
      double2 v;
      if (KIND == K_FWD_L)
        v = fv[(D.fvoff[t] + pp[k]) * nrhs + j];
      else if (KIND == K_FWD_L21)
        v = y[(dbeg + k) * nrhs + j];
      else if (KIND == K_BWD_U12)
        v = x[(int64_t)D.uidx[D.uoff[t] + k] * nrhs + j];
      else if (KIND == K_BWD_U)
        v = tt[(dbeg + k) * nrhs + j];
      else if (KIND == K_ROOT_L)
        v = fv[(D.fvoff[t] + O[rpp[k]]) * nrhs + j];
      else
        v = yr[k * nrhs + j];
      sin[kk * nsub + m] = v;
    }
    __syncthreads();
    const T *Mc = M + ic + ld * kc;
    for (int kk0 = 0; kk0 < kn; kk0 += kUnroll) {
      T araw[kUnroll];
#pragma unroll
      for (int u = 0; u < kUnroll; u++) araw[u] = Mc[ld * min(kk0 + u, kn - 1)];
#pragma unroll
      for (int u = 0; u < kUnroll; u++) {
        const int64_t k = kc + kk0 + u;
        double2 a = d2(araw[u]);
        bool use = kk0 + u < kn;
        if (KIND == K_FWD_L || KIND == K_ROOT_L) {
          if (k == i) a = make_double2(1.0, 0.0);
          use = use && k <= i;
        } else if (KIND == K_BWD_U || KIND == K_ROOT_U) {
          use = use && k >= i;
        }
        if (!use) a = make_double2(0.0, 0.0);
        const int kk = min(kk0 + u, kn - 1);
#pragma unroll
        for (int mm = 0; mm < MAXR; mm++) {
          if (mm >= my_n) break;
          const double2 b = sin[kk * nsub + ty + mm * ty_n];
          acc[mm].x += a.x * b.x - a.y * b.y;
          acc[mm].y += a.x * b.y + a.y * b.x;
        }
      }
    }
    __syncthreads();
  }
  if (i >= rows) return;
  double2 *pq = part + q * part_qstride + ((int64_t)tk.w * rows_wg + tx) * nrhs;
#pragma unroll
  for (int mm = 0; mm < MAXR; mm++) {
    if (mm >= my_n) break;
    pq[j0 + (ty + mm * ty_n) * rstep] = acc[mm];
  }
}

// Phase 1 on MFMA for FP32 factor blocks and at most 16 right-hand sides per launch row group (ty_n == 1): the
// same tasks, partial-sum slots and output as k_lu_apply_part.  Wave w takes rows 64 w .. 64 w + 63 of the task's
// 256 rows as four 16-row blocks; each 16 x 4 block of the factor is the first operand of v_mfma_f32_16x16x4f32
// (row i in lane i + 16 k), the staged inputs (4 columns x 16 right-hand sides, FP32, zero past nsub) the second;
// complex products by the three-product form (P1 = Ar.Xr, P2 = Ai.Xi, P3 = (Ar + Ai)(Xr + Xi); Re = P1 - P2,
// Im = P3 - P1 - P2).  Sums over a chunk are FP32 (the factor and the inner Krylov vectors are FP32); the slots are
// added in FP64 by k_lu_apply_sum.  The triangular blocks are masked per element (unit lower: k < i, diagonal 1;
// upper: k >= i).  NARROW (at most 8 right-hand sides): the second operand holds the real parts of the 8 inputs in
// columns 0-7 and their imaginary parts in columns 8-15, so two products P1 = Ar.[Xr|Xi], P2 = Ai.[Xr|Xi] give
// Re = P1[:, m] - P2[:, m + 8] and Im = P1[:, m + 8] + P2[:, m] (lanes m and m + 8 exchange in the epilogue): two
// MFMAs per block instead of three and 32 accumulators instead of 48.
template <int KIND, bool NARROW>
__global__ __launch_bounds__(256) void k_lu_apply_mfma(const int4 *__restrict__ tasks, LuDev D, RootDev Rt, int nrhs,
                                                       const float2 *__restrict__ Fall,
                                                       const float2 *__restrict__ Eall,
                                                       const int32_t *__restrict__ ppall,
                                                       const double2 *__restrict__ fvall,
                                                       const double2 *__restrict__ yall,
                                                       const double2 *__restrict__ tall,
                                                       const double2 *__restrict__ xall,
                                                       const double2 *__restrict__ yrall, float2 *__restrict__ part,
                                                       int64_t part_qstride, int kch) {
  constexpr int kKM = 128;         // columns staged per pass (one pass for the default chunk of 128 columns)
  __shared__ float2 sX[kKM][16];  // staged inputs: column k of the chunk, right-hand side m
  const int q = blockIdx.y;
  if (D.iact && !D.iact[q]) return;
  const int4 tk = tasks[blockIdx.x];
  int t, e;
  int64_t rows, cols, ld, dbeg, nd, nf;
  apply_geom<KIND>(tk, D, Rt, rows, cols, ld, t, e, dbeg, nd, nf);
  int j0 = 0, rstep = 1;
  if (KIND == K_ROOT_L || KIND == K_ROOT_U) {
    j0 = e;
    rstep = Rt.nexc;
  }
  const int nsub = (nrhs - j0 + rstep - 1) / rstep;
  const float2 *M;
  if (KIND == K_ROOT_L || KIND == K_ROOT_U) {
    M = Eall + ((int64_t)q * Rt.nexc + e) * Rt.estride;
  } else {
    const float2 *F = Fall + q * D.fstride + D.foff[t];
    M = KIND == K_FWD_L21 ? F + nd : KIND == K_BWD_U12 ? F + nf * nd : F;
  }
  const int64_t nrows_all = D.n * nrhs;
  const double2 *fv = fvall + q * D.fvstride;
  const double2 *y = yall + q * nrows_all, *tt = tall + q * nrows_all, *x = xall + q * nrows_all;
  const int32_t *pp = ppall + q * D.ppqs + dbeg;
  const int32_t *rpp = (KIND == K_ROOT_L) ? Rt.rpp + q * D.ppqs + (int64_t)e * Rt.no_max : nullptr;
  const int32_t *O = (KIND == K_ROOT_L || KIND == K_ROOT_U) ? Rt.O + (int64_t)e * Rt.no_max : nullptr;
  const double2 *yr = yrall + (int64_t)q * Rt.no_max * nrhs;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, li = lane & 15, lk = lane >> 4;
  const int64_t rbase = tk.y + 64 * w;  // first row of this wave
  const int64_t c_beg = tk.z, c_end = min((int64_t)tk.z + kch, cols);
  lu_f4 p1[4], p2[4], p3[NARROW ? 1 : 4];
#pragma unroll
  for (int rb = 0; rb < 4; rb++) {
    p1[rb] = lu_f4{0, 0, 0, 0};
    p2[rb] = lu_f4{0, 0, 0, 0};
    if (!NARROW) p3[rb] = lu_f4{0, 0, 0, 0};
  }
  // the rows this lane's first-operand values come from (clamped, so every load stays inside the block)
  int64_t arow[4];
#pragma unroll
  for (int rb = 0; rb < 4; rb++) arow[rb] = min(rbase + 16 * rb + li, rows - 1);
  constexpr int kGB = 2;
  // one group = kGB column blocks of 4 (4 kGB columns): 4 kGB factor values per lane, loaded before they are used
  auto load_group = [&](int g, float2 *buf, int64_t kc, int kn) {
#pragma unroll
    for (int u = 0; u < kGB; u++) {
      const int64_t kg = kc + min(4 * kGB * g + 4 * u + lk, kn - 1);
#pragma unroll
      for (int rb = 0; rb < 4; rb++) buf[4 * u + rb] = M[arow[rb] + ld * kg];
    }
  };
  auto mfma_group = [&](int g, const float2 *buf, int64_t kc, int kn) {
#pragma unroll
    for (int u = 0; u < kGB; u++) {
      const int kk = 4 * kGB * g + 4 * u + lk;  // this lane's column of the 16 x 4 block (and row of the 4 x 16 one)
      const int64_t kg = kc + min(kk, kn - 1);
      const float2 xv = sX[min(kk, kKM - 1)][NARROW ? (li & 7) : li];
      const float xs = xv.x + xv.y;
      const float xq = li < 8 ? xv.x : xv.y;  // NARROW: [Xr | Xi]
#pragma unroll
      for (int rb = 0; rb < 4; rb++) {
        const int64_t ig = rbase + 16 * rb + li;
        float2 a = buf[4 * u + rb];
        bool use = kk < kn;
        if (KIND == K_FWD_L || KIND == K_ROOT_L) {
          if (kg == ig) a = make_float2(1.f, 0.f);
          use = use && kg <= ig;
        } else if (KIND == K_BWD_U || KIND == K_ROOT_U) {
          use = use && kg >= ig;
        }
        if (!use) a = make_float2(0.f, 0.f);
        if (NARROW) {
          p1[rb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a.x, xq, p1[rb], 0, 0, 0);
          p2[rb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a.y, xq, p2[rb], 0, 0, 0);
        } else {
          p1[rb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a.x, xv.x, p1[rb], 0, 0, 0);
          p2[rb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a.y, xv.y, p2[rb], 0, 0, 0);
          p3[rb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a.x + a.y, xs, p3[rb], 0, 0, 0);
        }
      }
    }
  };
  for (int64_t kc = c_beg; kc < c_end; kc += kKM) {
    const int kn = (int)min((int64_t)kKM, c_end - kc);
    // the first factor group does not depend on the inputs: its loads go out before the staging
    float2 bufA[4 * kGB], bufB[4 * kGB];
    load_group(0, bufA, kc, kn);
    // stage the inputs: every load issued (clamped, unconditional) before the LDS stores; at most 8 right-hand
    // sides (the usual case) take 4 loads per thread, 9 to 16 take 8
    auto in_at = [&](int kk, int m) -> double2 {
      const int j = j0 + m * rstep;
      const int64_t k = kc + kk;
      if (KIND == K_FWD_L) return fv[(D.fvoff[t] + pp[k]) * nrhs + j];
      if (KIND == K_FWD_L21) return y[(dbeg + k) * nrhs + j];
      if (KIND == K_BWD_U12) return x[(int64_t)D.uidx[D.uoff[t] + k] * nrhs + j];
      if (KIND == K_BWD_U) return tt[(dbeg + k) * nrhs + j];
      if (KIND == K_ROOT_L) return fv[(D.fvoff[t] + O[rpp[k]]) * nrhs + j];
      return yr[k * nrhs + j];
    };
    // half h of the right-hand sides (m = 8 h .. 8 h + 7); the second half only when there are more than 8
    for (int h = 0; h < (nsub > 8 ? 2 : 1); h++) {
      constexpr int kSt = kKM * 8 / 256;
      double2 dv[kSt];
#pragma unroll
      for (int u = 0; u < kSt; u++) {
        const int s = threadIdx.x + 256 * u;
        dv[u] = in_at(min(s >> 3, kn - 1), min(8 * h + (s & 7), nsub - 1));
      }
#pragma unroll
      for (int u = 0; u < kSt; u++) {
        const int s = threadIdx.x + 256 * u;
        const int kk = s >> 3, m = 8 * h + (s & 7);
        sX[kk][m] = (kk < kn && m < nsub) ? make_float2((float)dv[u].x, (float)dv[u].y) : make_float2(0.f, 0.f);
        if (nsub <= 8) sX[kk][m + 8] = make_float2(0.f, 0.f);
      }
    }
    __syncthreads();
    // two register buffers: the next group's loads are in flight while the current group runs on MFMA
    const int ng = (kn + 4 * kGB - 1) / (4 * kGB);
    int g = 0;
    // (the scheduling barriers keep the compiler from sinking the next group's loads below the MFMAs)
    for (; g + 1 < ng; g += 2) {
      load_group(g + 1, bufB, kc, kn);
      __builtin_amdgcn_sched_barrier(0);
      mfma_group(g, bufA, kc, kn);
      __builtin_amdgcn_sched_barrier(0);
      if (g + 2 < ng) load_group(g + 2, bufA, kc, kn);
      __builtin_amdgcn_sched_barrier(0);
      mfma_group(g + 1, bufB, kc, kn);
      __builtin_amdgcn_sched_barrier(0);
    }
    if (g < ng) mfma_group(g, bufA, kc, kn);
    __syncthreads();
  }
  // D[i][m] of row block rb: lane l, register r: i = 4 (l / 16) + r, m = l % 16; the FP32 chunk sums are stored in
  // FP32 (k_lu_apply_sum adds them in FP64)
  float2 *pq = part + q * part_qstride + (int64_t)tk.w * 256 * nrhs;
  if (NARROW) {
    // lane m (< 8) holds P1[:, m], P2[:, m]; lane m + 8 holds P1[:, m + 8], P2[:, m + 8]
    float o1[4][4], o2[4][4];
#pragma unroll
    for (int rb = 0; rb < 4; rb++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        o1[rb][r] = __shfl_xor(p1[rb][r], 8, 16);
        o2[rb][r] = __shfl_xor(p2[rb][r], 8, 16);
      }
    if (li < 8 && li < nsub) {
      const int j = j0 + li * rstep;
#pragma unroll
      for (int rb = 0; rb < 4; rb++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int64_t ig = rbase + 16 * rb + 4 * lk + r;
          if (ig < rows) {
            const float re = p1[rb][r] - o2[rb][r], im = o1[rb][r] + p2[rb][r];
            pq[(ig - tk.y) * nrhs + j] = make_float2(re, im);
          }
        }
    }
  } else if (li < nsub) {
    const int j = j0 + li * rstep;
#pragma unroll
    for (int rb = 0; rb < 4; rb++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int64_t ig = rbase + 16 * rb + 4 * lk + r;
        if (ig < rows) {
          const float re = p1[rb][r] - p2[rb][r], im = p3[rb][r] - p1[rb][r] - p2[rb][r];
          pq[(ig - tk.y) * nrhs + j] = make_float2(re, im);
        }
      }
  }
}

// P: the partial-sum type (double2 from the vector kernels, float2 from the MFMA kernel); sums are formed in FP64.
template <int KIND, typename P>
__global__ __launch_bounds__(256) void k_lu_apply_sum(const int4 *__restrict__ tasks, LuDev D, RootDev Rt, int nrhs,
                                                      int ty_n, double2 *__restrict__ fvall,
                                                      double2 *__restrict__ yall, double2 *__restrict__ tall,
                                                      double2 *__restrict__ xall, double2 *__restrict__ yrall,
                                                      const P *__restrict__ part, int64_t part_qstride) {
  const int q = blockIdx.y;
  if (D.iact && !D.iact[q]) return;
  const int4 tk = tasks[blockIdx.x];
  const int rows_wg = 256 / ty_n;
  int t, e;
  int64_t rows, cols, ld, dbeg, nd, nf;
  apply_geom<KIND>(tk, D, Rt, rows, cols, ld, t, e, dbeg, nd, nf);
  int j0 = 0, rstep = 1;
  if (KIND == K_ROOT_L || KIND == K_ROOT_U) {
    j0 = e;
    rstep = Rt.nexc;
  }
  const int nsub = (nrhs - j0 + rstep - 1) / rstep;
  const int64_t nrows_all = D.n * nrhs;
  double2 *fv = fvall + q * D.fvstride;
  double2 *y = yall + q * nrows_all, *tt = tall + q * nrows_all, *x = xall + q * nrows_all;
  const int32_t *O = (KIND == K_ROOT_L || KIND == K_ROOT_U) ? Rt.O + (int64_t)e * Rt.no_max : nullptr;
  const P *pq = part + q * part_qstride;
  const int nrow = (int)min((int64_t)rows_wg, rows - tk.y);
  // one (row, right-hand side) pair per thread and pass, right-hand sides fastest, so consecutive threads read
  // consecutive partial sums of a slot; the slots are summed four at a time with independent loads
  for (int pi = threadIdx.x; pi < nrow * nsub; pi += 256) {
    const int tx = pi / nsub, m = pi % nsub;
    const int64_t i = tk.y + tx;
    const int j = j0 + m * rstep;
    const P *src = pq + ((int64_t)tk.w * rows_wg + tx) * nrhs + j;
    const int64_t cstride = (int64_t)rows_wg * nrhs;
    double2 a0 = make_double2(0.0, 0.0), a1 = a0, a2 = a0, a3 = a0;
    int c = 0;
    for (; c + 4 <= tk.z; c += 4) {
      const P v0 = src[(c + 0) * cstride], v1 = src[(c + 1) * cstride];
      const P v2 = src[(c + 2) * cstride], v3 = src[(c + 3) * cstride];
      a0.x += v0.x;
      a0.y += v0.y;
      a1.x += v1.x;
      a1.y += v1.y;
      a2.x += v2.x;
      a2.y += v2.y;
      a3.x += v3.x;
      a3.y += v3.y;
    }
    for (; c < tk.z; c++) {
      const P v = src[c * cstride];
      a0.x += v.x;
      a0.y += v.y;
    }
    const double2 acc = make_double2((a0.x + a1.x) + (a2.x + a3.x), (a0.y + a1.y) + (a2.y + a3.y));
    if (KIND == K_FWD_L) {
      y[(dbeg + i) * nrhs + j] = acc;
    } else if (KIND == K_FWD_L21) {
      double2 &d = fv[(D.fvoff[t] + nd + i) * nrhs + j];
      d.x -= acc.x;
      d.y -= acc.y;
    } else if (KIND == K_BWD_U12) {
      const double2 yv = y[(dbeg + i) * nrhs + j];
      tt[(dbeg + i) * nrhs + j] = make_double2(yv.x - acc.x, yv.y - acc.y);
    } else if (KIND == K_BWD_U) {
      x[(dbeg + i) * nrhs + j] = acc;
    } else if (KIND == K_ROOT_L) {
      yrall[(int64_t)q * Rt.no_max * nrhs + i * nrhs + j] = acc;
    } else {
      x[(dbeg + O[i]) * nrhs + j] = acc;
    }
  }
}

__global__ void k_lu_zero_rows(int64_t first, int64_t count, int nrhs, int64_t qstride, double2 *__restrict__ x,
                               const int32_t *__restrict__ iact) {
  const int q = blockIdx.y;
  if (iact && !iact[q]) return;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < count * nrhs;
       e += (int64_t)gridDim.x * blockDim.x)
    x[q * qstride + first * nrhs + e] = make_double2(0.0, 0.0);
}

inline int grid_for(int64_t n, int cap = 8192) {
  int64_t b = (n + 255) / 256;
  return (int)std::max<int64_t>(1, std::min<int64_t>(b, cap));
}

}  // namespace

// ---------------------------------------------------------------- host side
struct LuFactor::Impl {
  hipStream_t st = nullptr;
  // tree on the device
  DevBuf dbeg, uoff, foff, fvoff, uidx, umap, parent, kid_off, kids, perm, asm_pos, excl;
  int64_t fstride = 0, fvtot = 0;
  // schedule
  std::vector<std::vector<int32_t>> level_nodes;  // per level
  DevBuf d_ea;
  std::vector<int64_t> ea_off;
  int ea_cb = 16;
  // apply's forward gather: per level, (node, first row) init tasks; per (level, child slot), (child, first row)
  DevBuf d_fvi, d_fva;
  std::vector<int64_t> fvi_off, fva_off;
  // root
  int root = -1, nexc = 1, no_max = 0;
  std::vector<int32_t> no;
  DevBuf d_O, d_no;
  // blocked dense LU: fronts (tree nodes, then the port blocks), per-level step lists
  struct Step {
    int k0;
    int64_t panel_off, panel_n, lrows_off, lrows_n, laswp_off, laswp_n, trsm_off, trsm_n, gemm_off, gemm_n,
        gemm_tiles;
    double gemm_flops;  // real flops of the step's trailing updates, per matrix (8 per complex multiply-add)
    int64_t gemm_tiles128;  // 128 x 128 tiles of the same jobs (prefix in h_gprefix128)
    bool deep;              // the trailing update of an outer block (FP32 fronts may take the 128-tile kernel)
    bool deep_fit;          // at least 85% of the 128-tile area is matrix (the default rule for taking it)
    // symmetric factorization: tiles of the same jobs with triangular jobs cut to their lower tiles (prefixes in
    // h_gprefix_sym / h_gprefix128_sym), the real flops of those lower triangles (diagonal included), and the
    // mirror tasks run before the step (the first step of each outer block)
    int64_t gemm_tiles_sym, gemm_tiles128_sym;
    double gemm_flops_sym;
    int64_t mirror_off, mirror_n;
    int64_t lrows_sym_off, lrows_sym_n;  // rows from the end of the outer block on, for k_dense_lrows (sym)
  };
  struct DLevel {
    std::vector<Step> lu, inv;
    int64_t list_off = 0, list_n = 0;  // fronts of the level
    int64_t cols_off = 0, cols_n = 0;  // (front, 16-column chunk) tasks of the eye / put kernels
    int64_t xsize = 0;                 // inverse scratch elements per matrix
  };
  std::vector<DenseFront> dfr;
  std::vector<DLevel> dlev;    // tree levels, then one level of port blocks
  std::vector<int32_t> h_list;
  std::vector<int2> h_laswp, h_trsm_tasks, h_lrows, h_lrows_sym, h_cols;
  std::vector<TrsmJob> h_trsm;
  std::vector<GemmJob> h_gemm;
  std::vector<int64_t> h_gprefix, h_gprefix128, h_gprefix_sym, h_gprefix128_sym;
  std::vector<int4> h_mirror;
  DevBuf d_dfr, d_list, d_laswp, d_trsm_tasks, d_trsm, d_gemm, d_gprefix, d_gprefix128, d_gprefix_sym,
      d_gprefix128_sym, d_lrows, d_lrows_sym, d_cols, d_mirror;
  int64_t xmax = 1, ppqs = 1;
  // numeric storage
  DevBuf fronts, pp, sw, zp, E, X, Y;
  // apply storage and per-ty_n task lists
  DevBuf b, y, t, x, fv, yr;
  struct ApplyTasks {
    DevBuf d1, d2;
    std::vector<int64_t> off1[6], off2[6];  // per kind: task offsets per level (nlevels + 1)
    std::vector<int4> all1, all2;           // phase-1 (id, r0, c0, slot) and phase-2 (id, r0, chunks, slot0) tasks
    std::vector<double> elems[6];           // per kind and level: matrix entries one apply must read, per matrix
    int64_t max_slots = 1;                  // partial-sum slots of the largest launch
  };
  DevBuf part;
  DevBuf iact;  // per instance: any system active (apply)
  std::map<int, ApplyTasks> atasks;  // keyed by ty_n
  // kernel timing (OPEM_LU_TIMING set): HIP events on the stream around each trailing update, panel and apply
  // launch, drained when the stream is synchronised; class 0 trailing updates on the 64-tile kernel (work =
  // flops), 1 panels, 2 apply matrix products (work = matrix bytes read), 3 trailing updates on the 128-tile kernel
  bool timing = getenv("OPEM_LU_TIMING") != nullptr;
  // OPEM_LU_APPLY_MFMA=0 keeps the FP32-factor apply on the vector kernel
  bool apply_mfma = !(getenv("OPEM_LU_APPLY_MFMA") && getenv("OPEM_LU_APPLY_MFMA")[0] == '0');
  // columns per phase-1 chunk of the apply (the split-K depth; fewer chunks, fewer partial sums)
  int kch = [] {
    const char *e = getenv("OPEM_LU_KCH");
    const int v = e ? atoi(e) : kKCH;
    return v == 32 || v == 64 || v == 256 || v == 512 ? v : kKCH;
  }();
  // deep trailing updates of FP32 fronts: OPEM_LU_DEEP_GEMM=0 keeps them on the 64-tile kernel, 2 puts every one
  // on the 128-tile kernel, 1 (default) those whose jobs fill at least 85% of the 128-tile area
  int deep_gemm = getenv("OPEM_LU_DEEP_GEMM") ? std::max(0, std::min(2, atoi(getenv("OPEM_LU_DEEP_GEMM")))) : 1;
  // symmetric factorization of matrices the caller declares symmetric: OPEM_LU_SYM=0 never, 1 (default) FP32
  // factors, 2 FP32 and FP64 factors
  int sym_mode = getenv("OPEM_LU_SYM") ? std::max(0, std::min(2, atoi(getenv("OPEM_LU_SYM")))) : 1;
  // deep FP32 trailing updates with split-BF16 products (OPEM_LU_BF16=0: FP32 products)
  bool deep_bf16 = !(getenv("OPEM_LU_BF16") && getenv("OPEM_LU_BF16")[0] == '0');
  // the apply's level sweeps replayed from HIP graphs (OPEM_LU_GRAPH=0: launched one by one; always one by one
  // while kernel timing is on, whose events sit between the launches)
  bool use_graph = !(getenv("OPEM_LU_GRAPH") && getenv("OPEM_LU_GRAPH")[0] == '0');
  std::map<std::vector<int64_t>, hipGraphExec_t> graphs;
  void drop_graphs() {
    for (auto &g : graphs) (void)hipGraphExecDestroy(g.second);
    graphs.clear();
  }
  std::vector<hipEvent_t> evpool;
  size_t evused = 0;
  struct Pending {
    int cls;
    hipEvent_t a, b;
    double work;
  };
  std::vector<Pending> pending;
  double tsec[4] = {0, 0, 0, 0}, twork[4] = {0, 0, 0, 0}, tcount[4] = {0, 0, 0, 0};
  double factor_wall = 0, factors = 0, solves = 0;
  hipEvent_t ev() {
    if (evused == evpool.size()) {
      hipEvent_t e;
      HIPCHK(hipEventCreate(&e));
      evpool.push_back(e);
    }
    return evpool[evused++];
  }
  template <class F>
  void timed(int cls, double work, F launch) {
    if (!timing) {
      launch();
      return;
    }
    hipEvent_t a = ev(), b = ev();
    HIPCHK(hipEventRecord(a, st));
    launch();
    HIPCHK(hipEventRecord(b, st));
    pending.push_back({cls, a, b, work});
  }
  void drain() {  // the stream must be synchronised
    for (auto &p : pending) {
      float ms = 0;
      HIPCHK(hipEventElapsedTime(&ms, p.a, p.b));
      tsec[p.cls] += ms * 1e-3;
      twork[p.cls] += p.work;
      tcount[p.cls] += 1;
    }
    pending.clear();
    evused = 0;
  }
  void schedule_level(const std::vector<int32_t> &ids, DLevel &L);
  ~Impl() {
    drop_graphs();
    for (auto e : evpool) (void)hipEventDestroy(e);
  }
  LuDev dev(const LuTree &T, int nrhs) const {
    LuDev D;
    D.dbeg = dbeg.as<int64_t>();
    D.uoff = uoff.as<int64_t>();
    D.foff = foff.as<int64_t>();
    D.fvoff = fvoff.as<int64_t>();
    D.uidx = uidx.as<int32_t>();
    D.umap = umap.as<int32_t>();
    D.parent = parent.as<int32_t>();
    D.kid_off = kid_off.as<int32_t>();
    D.kids = kids.as<int32_t>();
    D.n = T.n;
    D.ppqs = ppqs;
    D.fstride = fstride;
    D.fvstride = fvtot * nrhs;
    D.iact = nullptr;
    return D;
  }
};

LuFactor::LuFactor() : d_(new Impl) {}
LuFactor::~LuFactor() { delete d_; }

size_t LuFactor::device_bytes_per_instance(int prec, int nrhs) const {
  if (!T_) return 0;
  const Impl &I = *d_;
  const LuTree &T = *T_;
  const double v = prec == 0 ? 16.0 : 8.0;
  double b = (double)I.fstride * v + (double)I.nexc * I.no_max * I.no_max * v + 2.0 * I.xmax * v + 2.0 * I.ppqs * 4;
  // apply: b, y, t, x, front vectors, root vectors, partial sums (slots of the largest launch, 64-row tiles bound)
  b += (4.0 * T.n + (double)I.fvtot + I.no_max) * nrhs * 16;
  double slots = 0;
  for (int l = 0; l < T.nlevels; l++) {
    double s2 = 0;
    for (int t : I.level_nodes[l]) s2 += (double)(T.nf(t) / 64 + 1) * (T.nf(t) / I.kch + 1);
    slots = std::max(slots, s2);
  }
  b += slots * 256 * nrhs * 16;
  return (size_t)b;
}

// Step lists of the blocked LU and of the two triangular inverses for the fronts `ids` (indices into dfr).
void LuFactor::Impl::schedule_level(const std::vector<int32_t> &ids, DLevel &L) {
  if (ids.empty()) return;
  int maxnd = 0;
  L.list_off = (int64_t)h_list.size();
  L.xsize = 0;
  for (int id : ids) {
    maxnd = std::max(maxnd, dfr[id].nd);
    dfr[id].x_off = L.xsize;
    L.xsize += (int64_t)dfr[id].nd * dfr[id].nd;
    h_list.push_back(id);
  }
  L.list_n = (int64_t)ids.size();
  xmax = std::max(xmax, L.xsize);
  L.cols_off = (int64_t)h_cols.size();
  for (int id : ids)
    for (int c0 = 0; c0 < dfr[id].nd; c0 += 16) h_cols.push_back(make_int2(id, c0));
  L.cols_n = (int64_t)h_cols.size() - L.cols_off;
  // a job joins step S: tile prefixes of the full and the triangular numbering, 64- and 128-tile kernels
  auto gemm_push = [&](GemmJob J, Step &S, bool tri) {
    J.tri = tri ? 1 : 0;
    h_gprefix.push_back(S.gemm_tiles);
    h_gprefix128.push_back(S.gemm_tiles128);
    h_gprefix_sym.push_back(S.gemm_tiles_sym);
    h_gprefix128_sym.push_back(S.gemm_tiles128_sym);
    h_gemm.push_back(J);
    const int64_t r64 = (J.m + kTile - 1) / kTile, c64 = (J.n + kTile - 1) / kTile;
    const int64_t r128 = (J.m + kBT - 1) / kBT, c128 = (J.n + kBT - 1) / kBT;
    S.gemm_tiles += r64 * c64;
    S.gemm_tiles128 += r128 * c128;
    S.gemm_tiles_sym += tri ? r64 * (r64 + 1) / 2 : r64 * c64;
    S.gemm_tiles128_sym += tri ? r128 * (r128 + 1) / 2 : r128 * c128;
    S.gemm_flops += 8.0 * J.m * J.n * J.k;
    S.gemm_flops_sym += tri ? 8.0 * J.k * (0.5 * J.m * (J.m + 1.0)) : 8.0 * J.m * J.n * J.k;
  };
  auto trsm_push = [&](const TrsmJob &J) {
    const int jid = (int)h_trsm.size();
    h_trsm.push_back(J);
    for (int c0 = 0; c0 < J.ncols; c0 += 256) h_trsm_tasks.push_back(make_int2(jid, c0));
  };
  // blocked LU in outer blocks of kOuter pivot columns: inner steps of kNB columns update only the rest of their
  // outer block; then the outer block's rows of U over the trailing columns (by inner blocks) and one trailing
  // update of depth kOuter, so the trailing matrix is read and written once per outer block
  auto empty_step = [&](int k0) {
    Step S{};
    S.k0 = k0;
    S.panel_off = (int64_t)h_list.size();
    S.lrows_off = (int64_t)h_lrows.size();
    S.laswp_off = (int64_t)h_laswp.size();
    S.trsm_off = (int64_t)h_trsm_tasks.size();
    S.gemm_off = (int64_t)h_gemm.size();
    S.mirror_off = (int64_t)h_mirror.size();
    S.lrows_sym_off = (int64_t)h_lrows_sym.size();
    return S;
  };
  auto close_step = [&](Step &S) {
    S.panel_n = (int64_t)h_list.size() - S.panel_off;
    S.lrows_n = (int64_t)h_lrows.size() - S.lrows_off;
    S.laswp_n = (int64_t)h_laswp.size() - S.laswp_off;
    S.trsm_n = (int64_t)h_trsm_tasks.size() - S.trsm_off;
    S.gemm_n = (int64_t)h_gemm.size() - S.gemm_off;
    L.lu.push_back(S);
  };
  for (int ob0 = 0; ob0 < maxnd; ob0 += kOuter) {
    for (int k0 = ob0; k0 < ob0 + kOuter && k0 < maxnd; k0 += kNB) {
      Step S = empty_step(k0);
      if (k0 == ob0)  // symmetric factorization: rows ob0 .. ob1-1 right of the diagonal from the columns below
        for (int id : ids) {
          const DenseFront &f = dfr[id];
          if (f.nd <= ob0) continue;
          const int ob1 = std::min(ob0 + kOuter, f.nd);
          for (int i0 = ob0; i0 < ob1; i0 += 64)
            for (int j0 = i0; j0 < f.nf; j0 += 64) h_mirror.push_back(make_int4(id, i0, j0, 0));
        }
      S.mirror_n = (int64_t)h_mirror.size() - S.mirror_off;
      S.panel_off = (int64_t)h_list.size();
      for (int id : ids)
        if (dfr[id].nd > k0) h_list.push_back(id);
      S.lrows_off = (int64_t)h_lrows.size();
      for (int id : ids)
        if (dfr[id].nd > k0)
          for (int r0 = 0; r0 < dfr[id].nf - dfr[id].nd; r0 += 256) h_lrows.push_back(make_int2(id, r0));
      S.lrows_sym_off = (int64_t)h_lrows_sym.size();
      for (int id : ids)
        if (dfr[id].nd > k0) {
          const int pend = std::min(dfr[id].nd, ob0 + kOuter);
          for (int r0 = 0; r0 < dfr[id].nf - pend; r0 += 256) h_lrows_sym.push_back(make_int2(id, r0));
        }
      S.lrows_sym_n = (int64_t)h_lrows_sym.size() - S.lrows_sym_off;
      S.laswp_off = (int64_t)h_laswp.size();
      for (int id : ids)
        if (dfr[id].nd > k0)
          for (int c0 = 0; c0 < dfr[id].nf; c0 += 256) h_laswp.push_back(make_int2(id, c0));
      S.trsm_off = (int64_t)h_trsm_tasks.size();
      S.gemm_off = (int64_t)h_gemm.size();
      for (int id : ids) {
        const DenseFront &f = dfr[id];
        if (f.nd <= k0) continue;
        const int nb = std::min(kNB, f.nd - k0), ob1 = std::min(ob0 + kOuter, f.nd), r = ob1 - k0 - nb;
        if (r <= 0) continue;
        TrsmJob T{};
        T.tsel = T.xsel = f.sel;
        T.nb = nb;
        T.ncols = r;
        T.toff = f.off + k0 + (int64_t)f.nf * k0;
        T.xoff = f.off + k0 + (int64_t)f.nf * (k0 + nb);
        T.ldt = T.ldx = f.nf;
        T.upper = 0;
        trsm_push(T);
        GemmJob G{};
        G.asel = G.bsel = G.csel = f.sel;
        G.m = f.nf - k0 - nb;
        G.n = r;
        G.k = nb;
        G.lda = G.ldb = G.ldc = f.nf;
        G.aoff = f.off + (k0 + nb) + (int64_t)f.nf * k0;
        G.boff = T.xoff;
        G.coff = f.off + (k0 + nb) + (int64_t)f.nf * (k0 + nb);
        gemm_push(G, S, false);
      }
      close_step(S);
    }
    // rows ob0 .. ob1 of U over the trailing columns ob1 .. nf: block forward substitution by inner blocks
    for (int k0 = ob0; k0 < ob0 + kOuter && k0 < maxnd; k0 += kNB) {
      Step S = empty_step(k0);
      for (int id : ids) {
        const DenseFront &f = dfr[id];
        if (f.nd <= k0) continue;
        const int nb = std::min(kNB, f.nd - k0), ob1 = std::min(ob0 + kOuter, f.nd), tc = f.nf - ob1;
        if (tc <= 0) continue;
        TrsmJob T{};
        T.tsel = T.xsel = f.sel;
        T.nb = nb;
        T.ncols = tc;
        T.toff = f.off + k0 + (int64_t)f.nf * k0;
        T.xoff = f.off + k0 + (int64_t)f.nf * ob1;
        T.ldt = T.ldx = f.nf;
        T.upper = 0;
        trsm_push(T);
        if (ob1 - k0 - nb > 0) {
          GemmJob G{};
          G.asel = G.bsel = G.csel = f.sel;
          G.m = ob1 - k0 - nb;
          G.n = tc;
          G.k = nb;
          G.lda = G.ldb = G.ldc = f.nf;
          G.aoff = f.off + (k0 + nb) + (int64_t)f.nf * k0;
          G.boff = T.xoff;
          G.coff = f.off + (k0 + nb) + (int64_t)f.nf * ob1;
          gemm_push(G, S, false);
        }
      }
      close_step(S);
    }
    // trailing update of depth ob1 - ob0
    {
      Step S = empty_step(ob0);
      S.deep = true;
      for (int id : ids) {
        const DenseFront &f = dfr[id];
        if (f.nd <= ob0) continue;
        const int ob1 = std::min(ob0 + kOuter, f.nd), mrem = f.nf - ob1;
        if (mrem <= 0) continue;
        GemmJob G{};
        G.asel = G.bsel = G.csel = f.sel;
        G.m = G.n = mrem;
        G.k = ob1 - ob0;
        G.lda = G.ldb = G.ldc = f.nf;
        G.aoff = f.off + ob1 + (int64_t)f.nf * ob0;
        G.boff = f.off + ob0 + (int64_t)f.nf * ob1;
        G.coff = f.off + ob1 + (int64_t)f.nf * ob1;
        gemm_push(G, S, true);
      }
      // the 128-tile kernel where its padding wastes little: at least 85% of its tile area is matrix
      double area = 0, area128 = 0;
      for (int64_t g = S.gemm_off; g < (int64_t)h_gemm.size(); g++) {
        const GemmJob &G = h_gemm[g];
        area += (double)G.m * G.n;
        area128 += (double)((G.m + kBT - 1) / kBT * kBT) * ((G.n + kBT - 1) / kBT * kBT);
      }
      S.deep_fit = area128 > 0 && area >= 0.85 * area128;
      close_step(S);
    }
  }
  // inverses: L^-1 block rows in ascending order, U^-1 block rows in descending order, in the same launches
  const int nsteps = (maxnd + kNB - 1) / kNB;
  for (int st_ = 0; st_ < nsteps; st_++) {
    Step S{};
    S.trsm_off = (int64_t)h_trsm_tasks.size();
    S.gemm_off = (int64_t)h_gemm.size();
    std::vector<GemmJob> gj;
    for (int id : ids) {
      const DenseFront &f = dfr[id];
      const int nblk = (f.nd + kNB - 1) / kNB;
      if (st_ >= nblk) continue;
      {  // L: block st_
        const int k0 = st_ * kNB, nb = std::min(kNB, f.nd - k0);
        TrsmJob T{};
        T.tsel = f.sel;
        T.xsel = 2;
        T.nb = nb;
        T.ncols = k0 + nb;
        T.toff = f.off + k0 + (int64_t)f.nf * k0;
        T.ldt = f.nf;
        T.xoff = f.x_off + k0;
        T.ldx = f.nd;
        T.upper = 0;
        trsm_push(T);
        if (f.nd - k0 - nb > 0) {
          GemmJob G{};
          G.asel = f.sel;
          G.bsel = G.csel = 2;
          G.m = f.nd - k0 - nb;
          G.n = k0 + nb;
          G.k = nb;
          G.lda = f.nf;
          G.ldb = G.ldc = f.nd;
          G.aoff = f.off + (k0 + nb) + (int64_t)f.nf * k0;
          G.boff = f.x_off + k0;
          G.coff = f.x_off + k0 + nb;
          gj.push_back(G);
        }
      }
      {  // U: block nblk - 1 - st_
        const int k0 = (nblk - 1 - st_) * kNB, nb = std::min(kNB, f.nd - k0);
        TrsmJob T{};
        T.tsel = f.sel;
        T.xsel = 3;
        T.nb = nb;
        T.ncols = f.nd - k0;
        T.toff = f.off + k0 + (int64_t)f.nf * k0;
        T.ldt = f.nf;
        T.xoff = f.x_off + k0 + (int64_t)f.nd * k0;
        T.ldx = f.nd;
        T.upper = 1;
        trsm_push(T);
        if (k0 > 0) {
          GemmJob G{};
          G.asel = f.sel;
          G.bsel = G.csel = 3;
          G.m = k0;
          G.n = f.nd - k0;
          G.k = nb;
          G.lda = f.nf;
          G.ldb = G.ldc = f.nd;
          G.aoff = f.off + (int64_t)f.nf * k0;
          G.boff = f.x_off + k0 + (int64_t)f.nd * k0;
          G.coff = f.x_off + (int64_t)f.nd * k0;
          gj.push_back(G);
        }
      }
    }
    for (auto &G : gj) gemm_push(G, S, false);
    S.trsm_n = (int64_t)h_trsm_tasks.size() - S.trsm_off;
    S.gemm_n = (int64_t)h_gemm.size() - S.gemm_off;
    L.inv.push_back(S);
  }
}

void LuFactor::setup(const LuTree &T, int nexc, const std::vector<int32_t> &root_group, hipStream_t st) {
  Impl &I = *d_;
  T_ = &T;
  I.st = st;
  const int nn = T.nnodes;
  I.dbeg.upload(T.dbeg, st);
  I.uoff.upload(T.uoff, st);
  I.foff.upload(T.foff, st);
  std::vector<int64_t> fvoff(nn + 1, 0);
  for (int t = 0; t < nn; t++) fvoff[t + 1] = fvoff[t] + T.nf(t);
  I.fvtot = fvoff[nn];
  I.fvoff.upload(fvoff, st);
  I.uidx.upload(T.uidx, st);
  I.umap.upload(T.umap, st);
  I.parent.upload(T.parent, st);
  std::vector<int32_t> kid_off(nn + 1, 0), kids;
  {
    std::vector<std::vector<int32_t>> kl(nn);
    for (int t = 0; t < nn; t++)
      if (T.parent[t] >= 0) kl[T.parent[t]].push_back(t);
    for (int t = 0; t < nn; t++) {
      kids.insert(kids.end(), kl[t].begin(), kl[t].end());
      kid_off[t + 1] = (int32_t)kids.size();
    }
  }
  I.kid_off.upload(kid_off, st);
  I.kids.upload(kids, st);
  I.perm.upload(T.perm, st);
  I.asm_pos.upload(T.asm_pos, st);
  {
    std::vector<int32_t> ex;  // rows outside the factored set (cleared in every apply's output)
    for (int64_t d = 0; d < T.ndof; d++)
      if (T.iperm[d] < 0) ex.push_back((int32_t)d);
    if ((int64_t)ex.size() != T.ndof - T.n) throw std::logic_error("lu: excluded rows do not match the tree");
    I.excl.upload(ex, st);
  }
  I.fstride = T.foff[nn];
  // root
  I.root = T.last_node;
  I.nexc = std::max(1, nexc);
  I.no.assign(I.nexc, 0);
  I.no_max = 1;
  if (I.root >= 0) {
    const int64_t ndr = T.nd(I.root);
    if ((int64_t)root_group.size() != ndr) throw std::invalid_argument("lu: root_group must cover the port rows");
    std::vector<std::vector<int32_t>> O(I.nexc);
    for (int64_t r = 0; r < ndr; r++) {
      if (root_group[r] < 0 || root_group[r] >= I.nexc) throw std::invalid_argument("lu: bad root group");
      for (int e = 0; e < I.nexc; e++)
        if (root_group[r] != e) O[e].push_back((int32_t)r);
    }
    for (int e = 0; e < I.nexc; e++) {
      I.no[e] = (int32_t)O[e].size();
      I.no_max = std::max(I.no_max, I.no[e]);
    }
    std::vector<int32_t> Of((size_t)I.nexc * I.no_max, 0);
    for (int e = 0; e < I.nexc; e++) std::copy(O[e].begin(), O[e].end(), Of.begin() + (size_t)e * I.no_max);
    I.d_O.upload(Of, st);
    I.d_no.upload(I.no, st);
  }
  I.ppqs = T.n + (int64_t)I.nexc * I.no_max;
  // schedule per level
  const int L = T.nlevels;
  I.level_nodes.assign(L, {});
  for (int t = 0; t < nn; t++) I.level_nodes[T.level[t]].push_back(t);
  // extend-add tasks: children of the level-l nodes, by slot
  const int MC = std::max(1, T.max_children);
  std::vector<int2> ea;
  I.ea_off.assign((size_t)L * MC + 1, 0);
  for (int l = 0; l < L; l++)
    for (int s = 0; s < MC; s++) {
      for (int p : I.level_nodes[l])
        for (int k = kid_off[p]; k < kid_off[p + 1]; k++) {
          const int c = kids[k];
          if (T.slot[c] != s) continue;
          for (int64_t b = 0; b < T.nu(c); b += I.ea_cb) ea.push_back(make_int2(c, (int)b));
        }
      I.ea_off[(size_t)l * MC + s + 1] = (int64_t)ea.size();
    }
  I.d_ea.upload(ea, st);
  {
    std::vector<int2> fvi, fva;
    I.fvi_off.assign((size_t)L + 1, 0);
    I.fva_off.assign((size_t)L * MC + 1, 0);
    for (int l = 0; l < L; l++) {
      for (int p : I.level_nodes[l])
        for (int64_t r = 0; r < T.nf(p); r += kFvRows) fvi.push_back(make_int2(p, (int)r));
      I.fvi_off[(size_t)l + 1] = (int64_t)fvi.size();
      for (int s = 0; s < MC; s++) {
        for (int p : I.level_nodes[l])
          for (int k = kid_off[p]; k < kid_off[p + 1]; k++) {
            const int c = kids[k];
            if (T.slot[c] != s) continue;
            for (int64_t r = 0; r < T.nu(c); r += kFvRows) fva.push_back(make_int2(c, (int)r));
          }
        I.fva_off[(size_t)l * MC + s + 1] = (int64_t)fva.size();
      }
    }
    I.d_fvi.upload(fvi, st);
    I.d_fva.upload(fva, st);
  }
  // dense fronts: tree nodes (the port root is only gathered), then one block per excitation
  I.dfr.assign(nn + I.nexc, DenseFront{});
  for (int t = 0; t < nn; t++) {
    DenseFront &f = I.dfr[t];
    f.sel = 0;
    f.nf = (int)T.nf(t);
    f.nd = (int)T.nd(t);
    f.off = T.foff[t];
    f.pp_off = T.dbeg[t];
  }
  const int64_t es = (int64_t)I.no_max * I.no_max;
  for (int e = 0; e < I.nexc; e++) {
    DenseFront &f = I.dfr[nn + e];
    f.sel = 1;
    f.nf = f.nd = I.root >= 0 ? I.no[e] : 0;
    f.off = (int64_t)e * es;
    f.pp_off = T.n + (int64_t)e * I.no_max;
  }
  I.h_list.clear();
  I.h_laswp.clear();
  I.h_trsm_tasks.clear();
  I.h_trsm.clear();
  I.h_gemm.clear();
  I.h_gprefix.clear();
  I.h_gprefix128.clear();
  I.h_gprefix_sym.clear();
  I.h_gprefix128_sym.clear();
  I.h_mirror.clear();
  I.h_lrows_sym.clear();
  I.h_lrows.clear();
  I.h_cols.clear();
  I.dlev.assign(L + 1, Impl::DLevel{});
  I.xmax = 1;
  for (int l = 0; l < L; l++) {
    std::vector<int32_t> ids;
    for (int t : I.level_nodes[l])
      if (t != I.root && T.nd(t) > 0) ids.push_back(t);
    I.schedule_level(ids, I.dlev[l]);
  }
  if (I.root >= 0) {
    std::vector<int32_t> ids;
    for (int e = 0; e < I.nexc; e++)
      if (I.no[e] > 0) ids.push_back(nn + e);
    I.schedule_level(ids, I.dlev[L]);
  }
  I.d_dfr.upload(I.dfr, st);
  I.d_list.upload(I.h_list, st);
  I.d_laswp.upload(I.h_laswp, st);
  I.d_trsm_tasks.upload(I.h_trsm_tasks, st);
  I.d_trsm.upload(I.h_trsm, st);
  I.d_gemm.upload(I.h_gemm, st);
  I.d_gprefix.upload(I.h_gprefix, st);
  I.d_gprefix128.upload(I.h_gprefix128, st);
  I.d_gprefix_sym.upload(I.h_gprefix_sym, st);
  I.d_gprefix128_sym.upload(I.h_gprefix128_sym, st);
  I.d_mirror.upload(I.h_mirror, st);
  I.d_lrows_sym.upload(I.h_lrows_sym, st);
  I.d_lrows.upload(I.h_lrows, st);
  I.d_cols.upload(I.h_cols, st);
  I.atasks.clear();
  I.drop_graphs();  // captured launch sequences belong to the previous tree
  HIPCHK(hipStreamSynchronize(st));
}

template <typename T>
static void dense_level(LuFactor::Impl &I, const LuFactor::Impl::DLevel &L, int ninst, const DenseBase &B, bool sym) {
  hipStream_t st = I.st;
  if (L.list_n == 0) return;
  const DenseFront *fr = I.d_dfr.as<DenseFront>();
  const int tri = sym ? 1 : 0;
  for (const auto &S : L.lu) {
    if (sym && S.mirror_n > 0)
      hipLaunchKernelGGL((k_dense_mirror<T>), dim3((unsigned)S.mirror_n, ninst), dim3(256), 0, st,
                         I.d_mirror.as<int4>() + S.mirror_off, fr, B, S.k0);
    if (S.panel_n > 0)
      I.timed(1, 0.0, [&] {
        if (sym)
          hipLaunchKernelGGL((k_dense_panel_blk<T>), dim3((unsigned)S.panel_n, ninst), dim3(kOuter), 0, st,
                             I.d_list.as<int32_t>() + S.panel_off, fr, B, S.k0, I.zp.as<unsigned long long>());
        else
          hipLaunchKernelGGL((k_dense_panel<T>), dim3((unsigned)S.panel_n, ninst), dim3(kPanelThreads), 0, st,
                             I.d_list.as<int32_t>() + S.panel_off, fr, B, S.k0, I.zp.as<unsigned long long>(), 0);
      });
    if (sym ? S.lrows_sym_n > 0 : S.lrows_n > 0)
      hipLaunchKernelGGL((k_dense_lrows<T>), dim3((unsigned)(sym ? S.lrows_sym_n : S.lrows_n), ninst), dim3(256), 0,
                         st, (sym ? I.d_lrows_sym : I.d_lrows).as<int2>() + (sym ? S.lrows_sym_off : S.lrows_off), fr,
                         B, S.k0, tri);
    if (S.laswp_n > 0)
      hipLaunchKernelGGL((k_dense_laswp<T>), dim3((unsigned)S.laswp_n, ninst), dim3(256), 0, st,
                         I.d_laswp.as<int2>() + S.laswp_off, fr, B, S.k0);
    if (S.trsm_n > 0)
      hipLaunchKernelGGL((k_dense_trsm<T>), dim3((unsigned)S.trsm_n, ninst), dim3(256), 0, st,
                         I.d_trsm_tasks.as<int2>() + S.trsm_off, I.d_trsm.as<TrsmJob>(), B);
    if (S.gemm_tiles > 0) {
      const double flops = sym ? S.gemm_flops_sym : S.gemm_flops;
      const int64_t t64 = sym ? S.gemm_tiles_sym : S.gemm_tiles, t128 = sym ? S.gemm_tiles128_sym : S.gemm_tiles128;
      const int64_t *p64 = (sym ? I.d_gprefix_sym : I.d_gprefix).as<int64_t>() + S.gemm_off;
      const int64_t *p128 = (sym ? I.d_gprefix128_sym : I.d_gprefix128).as<int64_t>() + S.gemm_off;
      if (std::is_same<T, float2>::value && S.deep && (I.deep_gemm == 2 || (I.deep_gemm == 1 && S.deep_fit)))
        I.timed(3, flops * ninst, [&] {
          if (I.deep_bf16)
            hipLaunchKernelGGL(k_dense_gemm_t128_bf16x3, dim3((unsigned)(t128 * ninst)), dim3(256), 0, st,
                               I.d_gemm.as<GemmJob>() + S.gemm_off, p128, (int)S.gemm_n, t128, ninst, B, tri);
          else
            hipLaunchKernelGGL(k_dense_gemm_t128, dim3((unsigned)(t128 * ninst)), dim3(256), 0, st,
                               I.d_gemm.as<GemmJob>() + S.gemm_off, p128, (int)S.gemm_n, t128, ninst, B, tri);
        });
      else
        I.timed(0, flops * ninst, [&] {
          hipLaunchKernelGGL((k_dense_gemm<T>), dim3((unsigned)t64, ninst), dim3(256), 0, st,
                             I.d_gemm.as<GemmJob>() + S.gemm_off, p64, (int)S.gemm_n, B, tri);
        });
    }
    LU_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL((k_dense_eye<T>), dim3((unsigned)L.cols_n, ninst), dim3(256), 0, st,
                     I.d_cols.as<int2>() + L.cols_off, fr, B);
  for (const auto &S : L.inv) {
    if (S.trsm_n > 0)
      hipLaunchKernelGGL((k_dense_trsm<T>), dim3((unsigned)S.trsm_n, ninst), dim3(256), 0, st,
                         I.d_trsm_tasks.as<int2>() + S.trsm_off, I.d_trsm.as<TrsmJob>(), B);
    if (S.gemm_tiles > 0)
      I.timed(0, S.gemm_flops * ninst, [&] {
        hipLaunchKernelGGL((k_dense_gemm<T>), dim3((unsigned)S.gemm_tiles, ninst), dim3(256), 0, st,
                           I.d_gemm.as<GemmJob>() + S.gemm_off, I.d_gprefix.as<int64_t>() + S.gemm_off,
                           (int)S.gemm_n, B, 0);
      });
    LU_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL((k_dense_put_inverses<T>), dim3((unsigned)L.cols_n, ninst), dim3(256), 0, st,
                     I.d_cols.as<int2>() + L.cols_off, fr, B);
  LU_LAUNCH_CHECK();
}

template <typename T>
static void factor_impl(LuFactor::Impl &I, const LuTree &Tr, int ninst, const double *vals, int64_t vstride,
                        LuFactorStats *stats, bool sym) {
  hipStream_t st = I.st;
  const int64_t fs = I.fstride;
  const int64_t es = (int64_t)I.no_max * I.no_max;
  I.fronts.alloc((size_t)ninst * fs * sizeof(T));
  I.pp.alloc((size_t)ninst * I.ppqs * 4);
  I.sw.alloc((size_t)ninst * I.ppqs * 4);
  I.E.alloc((size_t)ninst * I.nexc * es * sizeof(T));
  I.X.alloc((size_t)ninst * I.xmax * sizeof(T));
  I.Y.alloc((size_t)ninst * I.xmax * sizeof(T));
  I.zp.alloc(8);
  HIPCHK(hipMemsetAsync(I.zp.p, 0, 8, st));
  T *F = I.fronts.as<T>();
  DenseBase B{};
  B.p[0] = F;
  B.qs[0] = fs;
  B.p[1] = I.E.p;
  B.qs[1] = (int64_t)I.nexc * es;
  B.p[2] = I.X.p;
  B.qs[2] = I.xmax;
  B.p[3] = I.Y.p;
  B.qs[3] = I.xmax;
  B.pp = I.pp.as<int32_t>();
  B.sw = I.sw.as<int32_t>();
  B.ppqs = I.ppqs;
  LuDev D = I.dev(Tr, 1);
  hipLaunchKernelGGL((k_lu_zero<T>), dim3(grid_for((int64_t)ninst * fs, 65536)), dim3(256), 0, st,
                     (int64_t)ninst * fs, F);
  hipLaunchKernelGGL((k_lu_assemble<T>), dim3(grid_for((int64_t)Tr.asm_pos.size(), 16384), ninst), dim3(256), 0, st,
                     (int64_t)Tr.asm_pos.size(), I.asm_pos.as<int64_t>(), (const double2 *)vals, vstride, F, fs);
  LU_LAUNCH_CHECK();
  const int MC = std::max(1, Tr.max_children);
  for (int l = 0; l < Tr.nlevels; l++) {
    for (int s = 0; s < MC; s++) {
      const int64_t a = I.ea_off[(size_t)l * MC + s], b = I.ea_off[(size_t)l * MC + s + 1];
      if (b > a)
        hipLaunchKernelGGL((k_lu_extend_add<T>), dim3((unsigned)(b - a), ninst), dim3(256), 0, st,
                           I.d_ea.as<int2>() + a, D, F, I.ea_cb, sym ? 1 : 0);
    }
    LU_LAUNCH_CHECK();
    dense_level<T>(I, I.dlev[l], ninst, B, sym);
  }
  if (I.root >= 0) {
    const int r = I.root;
    const int nfr = (int)Tr.nd(r);
    T *E = I.E.as<T>();
    for (int e = 0; e < I.nexc; e++) {
      const int no = I.no[e];
      if (no == 0) continue;
      hipLaunchKernelGGL((k_lu_root_gather<T>), dim3(grid_for((int64_t)no * no, 16384), ninst), dim3(256), 0, st,
                         F + Tr.foff[r], fs, nfr, I.d_O.as<int32_t>() + (int64_t)e * I.no_max, no, E + (int64_t)e * es,
                         (int64_t)I.nexc * es, sym ? 1 : 0);
      LU_LAUNCH_CHECK();
    }
    dense_level<T>(I, I.dlev[Tr.nlevels], ninst, B, sym);
  }
  unsigned long long zp = 0;
  HIPCHK(hipMemcpyAsync(&zp, I.zp.p, 8, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (stats) {
    stats->zero_pivots = (int64_t)zp;
  }
}

void LuFactor::factor(int prec, int ninst, const double *vals, int64_t vals_stride, LuFactorStats *stats,
                      bool symmetric) {
  if (!T_) throw std::logic_error("lu: setup first");
  if (ninst < 1) throw std::invalid_argument("lu: ninst must be >= 1");
  const auto t0 = std::chrono::steady_clock::now();
  prec_ = prec;
  ninst_ = ninst;
  const bool sym = symmetric && (d_->sym_mode == 2 || (d_->sym_mode == 1 && prec != 0));
  sym_ = sym;
  if (prec == 0)
    factor_impl<double2>(*d_, *T_, ninst, vals, vals_stride, stats, sym);
  else
    factor_impl<float2>(*d_, *T_, ninst, vals, vals_stride, stats, sym);
  const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (stats) stats->seconds_factor = dt;
  d_->factor_wall += dt;
  d_->factors += ninst;
  if (d_->timing) d_->drain();  // factor_impl ends with a stream synchronisation
}

template <typename T, typename X>
static int solve_impl(LuFactor::Impl &I, const LuTree &Tr, int ninst, int nrhs, const X *v, int64_t ldv, X *z,
                      const int32_t *act) {
  if (nrhs > 8 * kMaxRhsThread) throw std::invalid_argument("lu: at most 64 right-hand sides per matrix");
  hipStream_t st = I.st;
  const int64_t n = Tr.n, S = (int64_t)ninst * nrhs;
  // right-hand sides per thread: 2, 4 or 8 (the template instance), rows per workgroup 256 / ty_n
  // (more right-hand sides per thread than 8 spill the accumulators to scratch); ty_n a power of two, so that row
  // groups tile the workgroup
  const int maxr = nrhs <= 2 ? 2 : nrhs <= 4 ? 4 : kMaxRhsThread;
  int ty_n = 1;
  while (ty_n * maxr < nrhs) ty_n *= 2;
  const int rows_wg = 256 / ty_n;
  const size_t vec = (size_t)ninst * std::max<int64_t>(n, 1) * nrhs * 16;
  I.b.alloc(vec);
  I.y.alloc(vec);
  I.t.alloc(vec);
  I.x.alloc(vec);
  I.fv.alloc((size_t)ninst * I.fvtot * nrhs * 16);
  I.yr.alloc((size_t)ninst * std::max(1, I.no_max) * nrhs * 16);
  // task lists for this row tiling
  auto it = I.atasks.find(ty_n);
  if (it == I.atasks.end()) {
    it = I.atasks.emplace(std::piecewise_construct, std::forward_as_tuple(ty_n), std::forward_as_tuple()).first;
    auto &A = it->second;
    const int L = Tr.nlevels;
    for (int kind = 0; kind < 6; kind++) {
      A.off1[kind].assign(L + 1, 0);
      A.off2[kind].assign(L + 1, 0);
    }
    // rows x cols of each kind's block and its triangle: 1 = lower with unit diagonal (k <= i), 2 = upper (k >= i)
    for (int kind = 0; kind < 6; kind++) A.elems[kind].assign(L, 0.0);
    double *cur_elems = nullptr;
    // OPEM_LU_APPLY_STATS (diagnostic): per kind, the phase-1 tasks, their mean fill of the rows x kch tile and the
    // share of tasks with at most 64 valid rows or 32 valid columns, printed once per row tiling
    const bool apply_stats = getenv("OPEM_LU_APPLY_STATS") != nullptr;
    double st_tasks[6] = {0}, st_fill[6] = {0}, st_small[6] = {0}, st_narrow[6] = {0};
    int cur_kind = 0;
    auto add_block = [&](int id, int64_t rows, int64_t cols, int tri, int &slot) {
      *cur_elems += tri == 0 ? (double)rows * cols : tri == 1 ? 0.5 * rows * (rows - 1) : 0.5 * rows * (rows + 1);
      for (int64_t r0 = 0; r0 < rows; r0 += rows_wg) {
        const int64_t r1 = std::min(rows, r0 + rows_wg);
        int64_t k0 = 0, k1 = cols;
        if (tri == 1) k1 = std::min(cols, r1);
        if (tri == 2) k0 = r0;
        const int slot0 = slot;
        int nch = 0;
        for (int64_t c0 = k0; c0 < k1; c0 += I.kch, nch++) {
          A.all1.push_back(make_int4(id, (int)r0, (int)c0, slot++));
          if (apply_stats) {  // diagnostic: valid rows and columns of each phase-1 task
            const double vr = (double)(r1 - r0), vc = (double)(std::min(k1, c0 + I.kch) - c0);
            st_tasks[cur_kind] += 1;
            st_fill[cur_kind] += vr * vc / ((double)rows_wg * I.kch);
            st_small[cur_kind] += vr <= 64;
            st_narrow[cur_kind] += vc <= 32;
          }
        }
        A.all2.push_back(make_int4(id, (int)r0, nch, slot0));
      }
    };
    for (int kind = 0; kind < 6; kind++) {
      for (int l = 0; l < L; l++) {
        A.off1[kind][l] = (int64_t)A.all1.size();
        A.off2[kind][l] = (int64_t)A.all2.size();
        cur_elems = &A.elems[kind][l];
        cur_kind = kind;
        int slot = 0;
        if (kind <= K_BWD_U) {
          for (int t : I.level_nodes[l]) {
            if (t == I.root) continue;
            const int64_t nd = Tr.nd(t), nu = Tr.nu(t);
            if (kind == K_FWD_L) add_block(t, nd, nd, 1, slot);
            if (kind == K_FWD_L21) add_block(t, nu, nd, 0, slot);
            if (kind == K_BWD_U12) add_block(t, nd, nu, 0, slot);
            if (kind == K_BWD_U) add_block(t, nd, nd, 2, slot);
          }
        } else if (I.root >= 0 && Tr.level[I.root] == l) {
          for (int e = 0; e < I.nexc; e++) add_block(e, I.no[e], I.no[e], kind == K_ROOT_L ? 1 : 2, slot);
        }
        A.max_slots = std::max<int64_t>(A.max_slots, slot);
      }
      A.off1[kind][L] = (int64_t)A.all1.size();
      A.off2[kind][L] = (int64_t)A.all2.size();
    }
    A.d1.upload(A.all1, st);
    A.d2.upload(A.all2, st);
    if (apply_stats)
      for (int kind = 0; kind < 6; kind++) {
        for (int l = 0; l < L; l++)
          if (A.off1[kind][l + 1] > A.off1[kind][l])
            fprintf(stderr, "lu apply kind %d level %d: %lld tasks, %.4e factor entries\n", kind, l,
                    (long long)(A.off1[kind][l + 1] - A.off1[kind][l]), A.elems[kind][l]);
        double el = 0;
        for (double v : A.elems[kind]) el += v;
        fprintf(stderr, "lu apply kind %d: %.0f tasks, %.3e factor entries, mean tile fill %.3f, rows<=64 %.3f, "
                "cols<=32 %.3f\n", kind, st_tasks[kind], el, st_fill[kind] / std::max(1.0, st_tasks[kind]),
                st_small[kind] / std::max(1.0, st_tasks[kind]), st_narrow[kind] / std::max(1.0, st_tasks[kind]));
      }
  }
  auto &A = it->second;
  const int64_t part_qstride = A.max_slots * rows_wg * nrhs;
  I.part.alloc((size_t)ninst * part_qstride * 16);
  double2 *part = I.part.as<double2>();
  // FP32 factors on the MFMA kernel: every right-hand side of a task in one MFMA tile, FP32 partial sums
  const bool mfma_path = std::is_same<T, float2>::value && ty_n == 1 && I.apply_mfma;
  float2 *part32 = I.part.as<float2>();
  LuDev D = I.dev(Tr, nrhs);
  // instances whose systems are all inactive (converged) are skipped by every apply kernel
  int nact = ninst;
  if (act) {
    I.iact.alloc((size_t)ninst * 4);
    hipLaunchKernelGGL(k_lu_inst_active, dim3((ninst + 63) / 64), dim3(64), 0, st, ninst, nrhs, act,
                       I.iact.as<int32_t>());
    LU_LAUNCH_CHECK();
    D.iact = I.iact.as<int32_t>();
    if (I.timing) {  // the bytes attributed to the timed launches count the active instances only
      std::vector<int32_t> ia(ninst);
      HIPCHK(hipMemcpyAsync(ia.data(), I.iact.p, (size_t)ninst * 4, hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
      nact = 0;
      for (int32_t a : ia) nact += a != 0;
    }
  }
  const int32_t *iactp = D.iact;
  RootDev Rt{};
  Rt.node = I.root;
  Rt.nexc = I.nexc;
  Rt.O = I.d_O.as<int32_t>();
  Rt.no = I.d_no.as<int32_t>();
  Rt.no_max = std::max(1, I.no_max);
  Rt.estride = (int64_t)Rt.no_max * Rt.no_max;
  Rt.rpp = I.pp.as<int32_t>() + Tr.n;
  const T *F = I.fronts.as<T>();
  const T *E = I.E.as<T>();
  const int32_t *pp = I.pp.as<int32_t>();
  double2 *b = I.b.as<double2>(), *y = I.y.as<double2>(), *tt = I.t.as<double2>(), *x = I.x.as<double2>(),
          *fv = I.fv.as<double2>(), *yr = I.yr.as<double2>();
  hipLaunchKernelGGL((k_lu_gather_rhs<X>), dim3(grid_for(n * nrhs), ninst), dim3(256), 0, st, n, nrhs,
                     I.perm.as<int32_t>(), v, ldv, b, iactp);
  LU_LAUNCH_CHECK();
  auto launch = [&](auto kind_c, int l) {
    constexpr int KIND = decltype(kind_c)::value;
    const int64_t a0 = A.off1[KIND][l], a1 = A.off1[KIND][l + 1];
    const int64_t b0 = A.off2[KIND][l], b1 = A.off2[KIND][l + 1];
    const size_t lds = (size_t)kKC * nrhs * sizeof(double2);
    if (a1 > a0) I.timed(2, A.elems[KIND][l] * sizeof(T) * nact, [&] {
      const dim3 g((unsigned)(a1 - a0), ninst);
      const int4 *tk = A.d1.as<int4>() + a0;
      if constexpr (std::is_same<T, float2>::value) {
        if (mfma_path) {
          if (nrhs <= 8)
            hipLaunchKernelGGL((k_lu_apply_mfma<KIND, true>), g, dim3(256), 0, st, tk, D, Rt, nrhs, F, E, pp, fv, y,
                               tt, x, yr, part32, part_qstride, I.kch);
          else
            hipLaunchKernelGGL((k_lu_apply_mfma<KIND, false>), g, dim3(256), 0, st, tk, D, Rt, nrhs, F, E, pp, fv, y,
                               tt, x, yr, part32, part_qstride, I.kch);
          return;
        }
      }
      if (maxr == 2)
        hipLaunchKernelGGL((k_lu_apply_part<T, KIND, 2>), g, dim3(256), lds, st, tk, D, Rt, nrhs, ty_n, F, E, pp, fv,
                           y, tt, x, yr, part, part_qstride, I.kch);
      else if (maxr == 4)
        hipLaunchKernelGGL((k_lu_apply_part<T, KIND, 4>), g, dim3(256), lds, st, tk, D, Rt, nrhs, ty_n, F, E, pp, fv,
                           y, tt, x, yr, part, part_qstride, I.kch);
      else
        hipLaunchKernelGGL((k_lu_apply_part<T, KIND, kMaxRhsThread>), g, dim3(256), lds, st, tk, D, Rt, nrhs, ty_n, F,
                           E, pp, fv, y, tt, x, yr, part, part_qstride, I.kch);
    });
    if (b1 > b0) {
      if (mfma_path)
        hipLaunchKernelGGL((k_lu_apply_sum<KIND, float2>), dim3((unsigned)(b1 - b0), ninst), dim3(256), 0, st,
                           A.d2.as<int4>() + b0, D, Rt, nrhs, ty_n, fv, y, tt, x, yr, part32, part_qstride);
      else
        hipLaunchKernelGGL((k_lu_apply_sum<KIND, double2>), dim3((unsigned)(b1 - b0), ninst), dim3(256), 0, st,
                           A.d2.as<int4>() + b0, D, Rt, nrhs, ty_n, fv, y, tt, x, yr, part, part_qstride);
    }
    LU_LAUNCH_CHECK();
  };
  using KL = std::integral_constant<int, K_FWD_L>;
  using KL21 = std::integral_constant<int, K_FWD_L21>;
  using KU12 = std::integral_constant<int, K_BWD_U12>;
  using KU = std::integral_constant<int, K_BWD_U>;
  using KRL = std::integral_constant<int, K_ROOT_L>;
  using KRU = std::integral_constant<int, K_ROOT_U>;
  const int L = Tr.nlevels;
  const int MC = std::max(1, Tr.max_children);
  // the level sweeps: a fixed launch sequence for a given tree, batch shape and set of buffers
  auto levels = [&] {
  for (int l = 0; l < L; l++) {
    {
      const int64_t a0 = I.fvi_off[l], a1 = I.fvi_off[l + 1];
      if (a1 > a0)
        hipLaunchKernelGGL(k_lu_fwd_init, dim3((unsigned)(a1 - a0), ninst), dim3(256), 0, st,
                           I.d_fvi.as<int2>() + a0, D, nrhs, b, fv);
      for (int s2 = 0; s2 < MC; s2++) {
        const int64_t c0 = I.fva_off[(size_t)l * MC + s2], c1 = I.fva_off[(size_t)l * MC + s2 + 1];
        if (c1 > c0)
          hipLaunchKernelGGL(k_lu_fwd_add, dim3((unsigned)(c1 - c0), ninst), dim3(256), 0, st,
                             I.d_fva.as<int2>() + c0, D, nrhs, fv);
      }
    }
    LU_LAUNCH_CHECK();
    launch(KL(), l);
    launch(KL21(), l);
    launch(KRL(), l);
  }
  if (I.root >= 0) {
    hipLaunchKernelGGL(k_lu_zero_rows, dim3(grid_for(Tr.nd(I.root) * nrhs), ninst), dim3(256), 0, st,
                       (int64_t)Tr.dbeg[I.root], (int64_t)Tr.nd(I.root), nrhs, n * nrhs, x, iactp);
    LU_LAUNCH_CHECK();
  }
  for (int l = L - 1; l >= 0; l--) {
    launch(KRU(), l);
    launch(KU12(), l);
    launch(KU(), l);
  }
  };
  if (I.use_graph && !I.timing) {
    // captured once per (precision, batch shape, buffers) and replayed: the kernels read the right-hand sides,
    // activity flags and factors through these fixed buffers, so only the launch sequence is frozen
    const std::vector<int64_t> key = {
        std::is_same<T, float2>::value ? 1 : 0, ty_n, ninst, nrhs, mfma_path ? 1 : 0, (int64_t)(uintptr_t)iactp,
        (int64_t)(uintptr_t)b, (int64_t)(uintptr_t)fv, (int64_t)(uintptr_t)y, (int64_t)(uintptr_t)tt,
        (int64_t)(uintptr_t)x, (int64_t)(uintptr_t)yr, (int64_t)(uintptr_t)part, (int64_t)(uintptr_t)F,
        (int64_t)(uintptr_t)E, (int64_t)(uintptr_t)pp};
    auto g = I.graphs.find(key);
    if (g == I.graphs.end()) {
      hipGraph_t graph = nullptr;
      HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      try {
        levels();
      } catch (...) {
        hipGraph_t aborted = nullptr;
        (void)hipStreamEndCapture(st, &aborted);
        if (aborted) (void)hipGraphDestroy(aborted);
        throw;
      }
      HIPCHK(hipStreamEndCapture(st, &graph));
      hipGraphExec_t exec = nullptr;
      const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      HIPCHK(ie);
      g = I.graphs.emplace(key, exec).first;
    }
    HIPCHK(hipGraphLaunch(g->second, st));
  } else {
    levels();
  }
  const int64_t nex = Tr.ndof - Tr.n;  // rows outside the factored set
  if (nex > 0)
    hipLaunchKernelGGL((k_lu_zero_list<X>), dim3(grid_for(nex * S, 65536)), dim3(256), 0, st, nex,
                       I.excl.as<int32_t>(), S, z);
  hipLaunchKernelGGL((k_lu_scatter<X>), dim3(grid_for(n * nrhs), ninst), dim3(256), 0, st, n, nrhs, S,
                     I.perm.as<int32_t>(), x, z, iactp);
  LU_LAUNCH_CHECK();
  return nact;
}

template <typename X>
void LuFactor::solve(int nrhs, const X *v, int64_t ldv, X *z, const int32_t *active) {
  if (!T_ || ninst_ < 1) throw std::logic_error("lu: factor first");
  const int nact = prec_ == 0 ? solve_impl<double2, X>(*d_, *T_, ninst_, nrhs, v, ldv, z, active)
                              : solve_impl<float2, X>(*d_, *T_, ninst_, nrhs, v, ldv, z, active);
  d_->solves += nact;
}
void LuFactor::kernel_stats(double *out, bool reset) {
  Impl &I = *d_;
  if (I.timing) {
    HIPCHK(hipStreamSynchronize(I.st));
    I.drain();
  }
  const double v[15] = {I.tsec[0], I.twork[0], I.tcount[0], I.tsec[1], I.tcount[1], I.tsec[2], I.twork[2], I.tcount[2],
                        I.factor_wall, I.factors, I.solves, I.timing ? 1.0 : 0.0, I.tsec[3], I.twork[3], I.tcount[3]};
  if (out)
    for (int k = 0; k < 15; k++) out[k] = v[k];
  if (!T_) return;
  if (reset) {
    for (int c = 0; c < 4; c++) I.tsec[c] = I.twork[c] = I.tcount[c] = 0;
    I.factor_wall = I.factors = I.solves = 0;
  }
}

template void LuFactor::solve<double2>(int, const double2 *, int64_t, double2 *, const int32_t *);
template void LuFactor::solve<float2>(int, const float2 *, int64_t, float2 *, const int32_t *);

}  // namespace opem
