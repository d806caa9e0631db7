// Numeric multifrontal sparse LU on the GPU: factorization and preconditioner apply (see lu.h).
//
// Factorization, level by level of the assembly tree (all fronts of a level are independent):
//   1. every front is cleared and receives its original matrix entries (one scatter over the pattern);
//   2. children's update blocks are added into their parents (one launch per child slot, so no two writers meet);
//   3. small fronts: one workgroup per (front, matrix) runs a right-looking LU in panels of 16 columns with row
//      pivoting inside the pivot block, then inverts the unit-lower and upper factors of the pivot block in place;
//      large fronts: the dense library (rocSOLVER getrf on the pivot rows, rocBLAS trsm + gemm for the Schur
//      complement, rocSOLVER trtri for the inverses), batched over the matrices.
// Apply (multi right-hand side, all levels): forward through the tree with front vectors (children's update
// parts added into the parent's), backward with the solved ancestors' values gathered through the update lists.
// Every dense step is y = M x with M a block of a factored front, so all threads of a level work at once.
#include "lu.h"

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <map>
#include <tuple>
#include <stdexcept>
#include <type_traits>

namespace opem {

#define ROCBLAS_CHK(x)                                                                          \
  do {                                                                                          \
    rocblas_status s_ = (x);                                                                    \
    if (s_ != rocblas_status_success)                                                           \
      throw std::runtime_error(std::string(#x) + ": " + rocblas_status_to_string(s_));          \
  } while (0)

#define LU_LAUNCH_CHECK()                                                                        \
  do {                                                                                           \
    hipError_t e_ = hipGetLastError();                                                           \
    if (e_ != hipSuccess) throw std::runtime_error(std::string("lu kernel launch: ") + hipGetErrorString(e_)); \
  } while (0)

namespace {

constexpr int kNB = 16;          // panel width of the small-front LU
constexpr int kSmallFront = 1536;  // fronts up to this order use the one-workgroup kernel
constexpr int kMaxRhsThread = 16;  // right-hand sides per thread in the apply kernels
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
  int64_t n, fstride, fvstride;
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
template <typename T>
__global__ __launch_bounds__(256) void k_lu_extend_add(const int2 *__restrict__ tasks, LuDev D, T *__restrict__ F,
                                                       int cb) {
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
    for (int64_t a = threadIdx.x; a < nuc; a += blockDim.x) {
      T v = Fp[um[a] + pc];
      const T s = src[a];
      v.x += s.x;
      v.y += s.y;
      Fp[um[a] + pc] = v;
    }
  }
}

// Partial LU of one front per workgroup (see the file comment).  pp[k] (k < nd): the front row that ended at pivot
// position k.  Exactly zero pivots are replaced by a tiny value and counted.
template <typename T>
__global__ __launch_bounds__(256) void k_lu_front_small(const int32_t *__restrict__ nodes, LuDev D,
                                                        T *__restrict__ Fall, int32_t *__restrict__ ppall,
                                                        unsigned long long *__restrict__ zero_pivots) {
  using R = decltype(T().x);
  const int t = nodes[blockIdx.x];
  const int q = blockIdx.y;
  const int nd = (int)(D.dbeg[t + 1] - D.dbeg[t]);
  const int nf = nd + (int)(D.uoff[t + 1] - D.uoff[t]);
  T *F = Fall + q * D.fstride + D.foff[t];
  int32_t *pp = ppall + q * D.n + D.dbeg[t];
  const int tid = threadIdx.x;
  __shared__ double red_v[256];
  __shared__ int red_i[256];
  __shared__ T Lkk[kNB * kNB];
  __shared__ T colbuf[kSmallFront];
  for (int k = tid; k < nd; k += 256) pp[k] = k;
  __syncthreads();
  for (int k0 = 0; k0 < nd; k0 += kNB) {
    const int nb = min(kNB, nd - k0);
    // ---- panel: columns k0 .. k0+nb-1
    for (int j = k0; j < k0 + nb; j++) {
      double best = -1.0;
      int bi = j;
      for (int i = j + tid; i < nd; i += 256) {
        const double v = cabs1(F[i + (int64_t)nf * j]);
        if (v > best) {
          best = v;
          bi = i;
        }
      }
      red_v[tid] = best;
      red_i[tid] = bi;
      __syncthreads();
      for (int s = 128; s > 0; s >>= 1) {
        if (tid < s) {
          const double v2 = red_v[tid + s];
          const int i2 = red_i[tid + s];
          if (v2 > red_v[tid] || (v2 == red_v[tid] && i2 < red_i[tid])) {
            red_v[tid] = v2;
            red_i[tid] = i2;
          }
        }
        __syncthreads();
      }
      const int piv = red_i[0];
      const bool zero = !(red_v[0] > 0.0);
      __syncthreads();
      if (piv != j) {
        for (int c = tid; c < nf; c += 256) {
          const T a = F[j + (int64_t)nf * c];
          F[j + (int64_t)nf * c] = F[piv + (int64_t)nf * c];
          F[piv + (int64_t)nf * c] = a;
        }
        if (tid == 0) {
          const int32_t a = pp[j];
          pp[j] = pp[piv];
          pp[piv] = a;
        }
      }
      __syncthreads();
      if (zero && tid == 0) {
        T tiny;
        tiny.x = std::is_same<R, float>::value ? (R)1e-30 : (R)1e-290;
        tiny.y = 0;
        F[j + (int64_t)nf * j] = tiny;
        atomicAdd(zero_pivots, 1ull);
      }
      __syncthreads();
      const T inv = cinv_t(F[j + (int64_t)nf * j]);
      for (int i = j + 1 + tid; i < nf; i += 256) F[i + (int64_t)nf * j] = cmul_t(F[i + (int64_t)nf * j], inv);
      __syncthreads();
      // rank-1 update of the remaining panel columns
      const int ncol = k0 + nb - (j + 1);
      if (ncol > 0) {
        for (int i = j + 1 + tid; i < nf; i += 256) {
          const T l = F[i + (int64_t)nf * j];
          for (int c = j + 1; c < k0 + nb; c++) {
            const T u = F[j + (int64_t)nf * c];
            F[i + (int64_t)nf * c] = csub_t(F[i + (int64_t)nf * c], cmul_t(l, u));
          }
        }
      }
      __syncthreads();
    }
    // ---- block row of U: rows k0 .. k0+nb-1, columns k0+nb .. nf-1  (unit lower L_kk solve)
    for (int e = tid; e < nb * nb; e += 256) Lkk[e] = F[(k0 + e % nb) + (int64_t)nf * (k0 + e / nb)];
    __syncthreads();
    for (int c = k0 + nb + tid; c < nf; c += 256) {
      T x[kNB];
#pragma unroll
      for (int r = 0; r < kNB; r++)
        if (r < nb) x[r] = F[k0 + r + (int64_t)nf * c];
#pragma unroll
      for (int r = 1; r < kNB; r++)
        if (r < nb) {
#pragma unroll
          for (int s = 0; s < kNB; s++)
            if (s < r) x[r] = csub_t(x[r], cmul_t(Lkk[r + nb * s], x[s]));
        }
#pragma unroll
      for (int r = 1; r < kNB; r++)
        if (r < nb) F[k0 + r + (int64_t)nf * c] = x[r];
    }
    __syncthreads();
    // ---- trailing update: rows and columns k0+nb .. nf-1
    const int r1 = k0 + nb;
    if (r1 < nf) {
      const int tx = tid & 63, ty = tid >> 6;
      for (int i0 = r1; i0 < nf; i0 += 64) {
        const int i = i0 + tx;
        T l[kNB];
        if (i < nf) {
#pragma unroll
          for (int s = 0; s < kNB; s++)
            if (s < nb) l[s] = F[i + (int64_t)nf * (k0 + s)];
        }
        for (int c = r1 + ty; c < nf; c += 4) {
          if (i < nf) {
            T acc = F[i + (int64_t)nf * c];
            const T *u = F + k0 + (int64_t)nf * c;
#pragma unroll
            for (int s = 0; s < kNB; s++)
              if (s < nb) acc = csub_t(acc, cmul_t(l[s], u[s]));
            F[i + (int64_t)nf * c] = acc;
          }
        }
      }
    }
    __syncthreads();
  }
  // ---- inverse of U11 in place (upper, with diagonal)
  for (int j = 0; j < nd; j++) {
    for (int k = tid; k <= j; k += 256) colbuf[k] = F[k + (int64_t)nf * j];
    __syncthreads();
    const T d = cinv_t(colbuf[j]);
    T nd_;
    nd_.x = -d.x;
    nd_.y = -d.y;
    for (int i = tid; i < j; i += 256) {
      double2 acc = make_double2(0.0, 0.0);
      for (int k = i; k < j; k++) {
        const double2 a = d2(F[i + (int64_t)nf * k]), b = d2(colbuf[k]);
        acc.x += a.x * b.x - a.y * b.y;
        acc.y += a.x * b.y + a.y * b.x;
      }
      F[i + (int64_t)nf * j] = cmul_t(nd_, cvt<T>(acc));
    }
    if (tid == 0) F[j + (int64_t)nf * j] = d;
    __syncthreads();
  }
  // ---- inverse of the unit-lower L11 in place (strictly lower part)
  for (int j = nd - 2; j >= 0; j--) {
    for (int k = j + 1 + tid; k < nd; k += 256) colbuf[k] = F[k + (int64_t)nf * j];
    __syncthreads();
    for (int i = j + 1 + tid; i < nd; i += 256) {
      double2 acc = d2(colbuf[i]);
      for (int k = j + 1; k < i; k++) {
        const double2 a = d2(F[i + (int64_t)nf * k]), b = d2(colbuf[k]);
        acc.x += a.x * b.x - a.y * b.y;
        acc.y += a.x * b.y + a.y * b.x;
      }
      F[i + (int64_t)nf * j] = cvt<T>(make_double2(-acc.x, -acc.y));
    }
    __syncthreads();
  }
}

// LAPACK-style row interchanges (1-based, relative to the block) -> the permutation pp[k] = block row at pivot k.
// One workgroup per (block, matrix); the swaps run in LDS on one thread.
__global__ void k_lu_ipiv_to_perm(const int64_t *__restrict__ off, const int32_t *__restrict__ len,
                                  const int32_t *__restrict__ ipiv, int32_t *__restrict__ pp, int64_t qstride) {
  extern __shared__ int32_t sp[];
  const int b = blockIdx.x, q = blockIdx.y;
  const int n = len[b];
  const int32_t *ip = ipiv + q * qstride + off[b];
  int32_t *out = pp + q * qstride + off[b];
  for (int k = threadIdx.x; k < n; k += blockDim.x) sp[k] = k;
  __syncthreads();
  if (threadIdx.x == 0)
    for (int k = 0; k < n; k++) {
      const int s = ip[k] - 1;
      if (s != k && s >= 0 && s < n) {
        const int32_t a = sp[k];
        sp[k] = sp[s];
        sp[s] = a;
      }
    }
  __syncthreads();
  for (int k = threadIdx.x; k < n; k += blockDim.x) out[k] = sp[k];
}

// Excitation e's port block: E[i + no*k] = S[O[i] + nfr * O[k]], S the root front (the ports' Schur complement).
template <typename T>
__global__ void k_lu_root_gather(const T *__restrict__ S, int64_t sstride, int nfr, const int32_t *__restrict__ O,
                                 int no, T *__restrict__ E, int64_t estride) {
  const int q = blockIdx.y;
  const int64_t tot = (int64_t)no * no;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(e % no), k = (int)(e / no);
    E[q * estride + e] = S[q * sstride + O[i] + (int64_t)nfr * O[k]];
  }
}

// ---------------------------------------------------------------- apply kernels
// Work vectors are double2 [matrix][row][nrhs].

template <typename X>
__global__ void k_lu_gather_rhs(int64_t n, int nrhs, const int32_t *__restrict__ perm, const X *__restrict__ v,
                                int64_t ldv, double2 *__restrict__ b) {
  const int q = blockIdx.y;
  const int64_t tot = n * nrhs;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t pos = e / nrhs;
    const int j = (int)(e % nrhs);
    b[q * tot + e] = d2(v[((int64_t)q * nrhs + j) * ldv + perm[pos]]);
  }
}

template <typename X>
__global__ void k_lu_scatter(int64_t n, int nrhs, int64_t S, const int32_t *__restrict__ perm,
                             const double2 *__restrict__ x, X *__restrict__ z) {
  const int q = blockIdx.y;
  const int64_t tot = n * nrhs;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t pos = e / nrhs;
    const int j = (int)(e % nrhs);
    z[(int64_t)perm[pos] * S + (int64_t)q * nrhs + j] = cvt<X>(x[q * tot + e]);
  }
}

// Front vectors of the nodes of one level: pivot rows from b, update rows zero, then each child's update part
// added at its umap rows (children one after another inside the workgroup).
__global__ __launch_bounds__(256) void k_lu_fwd_gather(const int32_t *__restrict__ nodes, LuDev D, int nrhs,
                                                       const double2 *__restrict__ ball, double2 *__restrict__ fvall) {
  const int t = nodes[blockIdx.x];
  const int q = blockIdx.y;
  const int64_t nd = D.dbeg[t + 1] - D.dbeg[t], nf = nd + (D.uoff[t + 1] - D.uoff[t]);
  double2 *fv = fvall + q * D.fvstride + D.fvoff[t] * nrhs;
  const double2 *b = ball + q * D.n * nrhs + D.dbeg[t] * nrhs;
  for (int64_t e = threadIdx.x; e < nf * nrhs; e += blockDim.x) fv[e] = e < nd * nrhs ? b[e] : make_double2(0.0, 0.0);
  __syncthreads();
  for (int k = D.kid_off[t]; k < D.kid_off[t + 1]; k++) {
    const int c = D.kids[k];
    const int64_t ndc = D.dbeg[c + 1] - D.dbeg[c], nuc = D.uoff[c + 1] - D.uoff[c];
    const double2 *fc = fvall + q * D.fvstride + (D.fvoff[c] + ndc) * nrhs;
    const int32_t *um = D.umap + D.uoff[c];
    for (int64_t e = threadIdx.x; e < nuc * nrhs; e += blockDim.x) {
      const int64_t a = e / nrhs;
      const int j = (int)(e % nrhs);
      double2 &dst = fv[(int64_t)um[a] * nrhs + j];
      dst.x += fc[e].x;
      dst.y += fc[e].y;
    }
    __syncthreads();
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
  const int32_t *rpp;    // [matrix][nexc][no_max] pivot permutation of the excitation blocks
};

// One task = (node or excitation, first row).  Each thread owns one row of the task and up to kMaxRhsThread
// right-hand sides (rhs j = j0 + rstep * m); the inputs are staged kKC rows at a time in LDS.
template <typename T, int KIND>
__global__ __launch_bounds__(256) void k_lu_apply(const int2 *__restrict__ tasks, LuDev D, RootDev Rt, int nrhs,
                                                  int ty_n, const T *__restrict__ Fall, const T *__restrict__ Eall,
                                                  const int32_t *__restrict__ ppall, double2 *__restrict__ fvall,
                                                  double2 *__restrict__ yall, double2 *__restrict__ tall,
                                                  double2 *__restrict__ xall, double2 *__restrict__ yrall) {
  __shared__ double2 sin[kKC * kMaxRhsThread * 4];
  const int2 tk = tasks[blockIdx.x];
  const int q = blockIdx.y;
  const int rows_wg = 256 / ty_n;
  const int tx = threadIdx.x % rows_wg, ty = threadIdx.x / rows_wg;
  // rhs subset of this task
  int j0 = 0, rstep = 1;
  if (KIND == K_ROOT_L || KIND == K_ROOT_U) {
    j0 = tk.x;
    rstep = Rt.nexc;
  }
  const int nsub = (nrhs - j0 + rstep - 1) / rstep;  // rhs in the subset
  // this thread's rhs: subset indices ty, ty + ty_n, ...
  const int my_n = nsub > ty ? (nsub - ty + ty_n - 1) / ty_n : 0;
  const int64_t nrows_all = D.n * nrhs;
  double2 *fv = fvall + q * D.fvstride;
  double2 *y = yall + q * nrows_all, *tt = tall + q * nrows_all, *x = xall + q * nrows_all;
  int t = -1, e = -1;
  int64_t nd = 0, nu = 0, nf = 0, dbeg = 0;
  const T *M = nullptr;
  int64_t ld = 0, rows = 0, cols = 0;
  if (KIND == K_ROOT_L || KIND == K_ROOT_U) {
    e = tk.x;
    t = Rt.node;
    nd = D.dbeg[t + 1] - D.dbeg[t];
    dbeg = D.dbeg[t];
    rows = cols = Rt.no[e];
    ld = rows;
    M = Eall + ((int64_t)q * Rt.nexc + e) * Rt.estride;
  } else {
    t = tk.x;
    dbeg = D.dbeg[t];
    nd = D.dbeg[t + 1] - dbeg;
    nu = D.uoff[t + 1] - D.uoff[t];
    nf = nd + nu;
    ld = nf;
    const T *F = Fall + q * D.fstride + D.foff[t];
    if (KIND == K_FWD_L || KIND == K_BWD_U) {
      M = F;
      rows = cols = nd;
    } else if (KIND == K_FWD_L21) {
      M = F + nd;
      rows = nu;
      cols = nd;
    } else {
      M = F + nf * nd;
      rows = nd;
      cols = nu;
    }
  }
  const int64_t i = tk.y + tx;
  double2 acc[kMaxRhsThread];
#pragma unroll
  for (int m = 0; m < kMaxRhsThread; m++) acc[m] = make_double2(0.0, 0.0);
  // column range touched by this task's rows
  int64_t k_lo = 0, k_hi = cols;
  if (KIND == K_FWD_L || KIND == K_ROOT_L) k_hi = min((int64_t)tk.y + rows_wg, rows);  // k <= i
  if (KIND == K_BWD_U || KIND == K_ROOT_U) k_lo = tk.y;                                 // k >= i
  const int32_t *pp = ppall + q * D.n + dbeg;
  const int32_t *rpp = (KIND == K_ROOT_L) ? Rt.rpp + ((int64_t)q * Rt.nexc + e) * Rt.no_max : nullptr;
  const int32_t *O = (KIND == K_ROOT_L || KIND == K_ROOT_U) ? Rt.O + (int64_t)e * Rt.no_max : nullptr;
  const double2 *yr = yrall + (int64_t)q * Rt.no_max * nrhs;
  for (int64_t kc = k_lo; kc < k_hi; kc += kKC) {
    const int kn = (int)min((int64_t)kKC, k_hi - kc);
    // stage inputs in[k][m] for the subset rhs (m < nsub)
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
    if (i < rows) {
      for (int kk = 0; kk < kn; kk++) {
        const int64_t k = kc + kk;
        double2 a;
        if (KIND == K_FWD_L || KIND == K_ROOT_L) {
          if (k > i) break;
          a = k == i ? make_double2(1.0, 0.0) : d2(M[i + ld * k]);
        } else if (KIND == K_BWD_U || KIND == K_ROOT_U) {
          if (k < i) continue;
          a = d2(M[i + ld * k]);
        } else {
          a = d2(M[i + ld * k]);
        }
#pragma unroll
        for (int mm = 0; mm < kMaxRhsThread; mm++) {
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
#pragma unroll
  for (int mm = 0; mm < kMaxRhsThread; mm++) {
    if (mm >= my_n) break;
    const int j = j0 + (ty + mm * ty_n) * rstep;
    if (KIND == K_FWD_L) {
      y[(dbeg + i) * nrhs + j] = acc[mm];
    } else if (KIND == K_FWD_L21) {
      double2 &d = fv[(D.fvoff[t] + nd + i) * nrhs + j];
      d.x -= acc[mm].x;
      d.y -= acc[mm].y;
    } else if (KIND == K_BWD_U12) {
      const double2 yv = y[(dbeg + i) * nrhs + j];
      tt[(dbeg + i) * nrhs + j] = make_double2(yv.x - acc[mm].x, yv.y - acc[mm].y);
    } else if (KIND == K_BWD_U) {
      x[(dbeg + i) * nrhs + j] = acc[mm];
    } else if (KIND == K_ROOT_L) {
      yrall[(int64_t)q * Rt.no_max * nrhs + i * nrhs + j] = acc[mm];
    } else {
      x[(dbeg + O[i]) * nrhs + j] = acc[mm];
    }
  }
}

__global__ void k_lu_zero_rows(int64_t first, int64_t count, int nrhs, int64_t qstride, double2 *__restrict__ x) {
  const int q = blockIdx.y;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < count * nrhs;
       e += (int64_t)gridDim.x * blockDim.x)
    x[q * qstride + first * nrhs + e] = make_double2(0.0, 0.0);
}

inline int grid_for(int64_t n, int cap = 8192) {
  int64_t b = (n + 255) / 256;
  return (int)std::max<int64_t>(1, std::min<int64_t>(b, cap));
}

// ---------------------------------------------------------------- dense library dispatch
template <typename T>
struct Lib;
template <>
struct Lib<double2> {
  using C = rocblas_double_complex;
  static rocblas_status getrf(rocblas_handle h, int m, int n, double2 *A, int lda, rocblas_stride sA, int *ip,
                              rocblas_stride sP, int *info, int bc) {
    return rocsolver_zgetrf_strided_batched(h, m, n, (C *)A, lda, sA, ip, sP, info, bc);
  }
  static rocblas_status trsm_ru(rocblas_handle h, int m, int n, const double2 *A, int lda, rocblas_stride sA,
                                double2 *B, int ldb, rocblas_stride sB, int bc) {
    const C one(1.0, 0.0);
    return rocblas_ztrsm_strided_batched(h, rocblas_side_right, rocblas_fill_upper, rocblas_operation_none,
                                         rocblas_diagonal_non_unit, m, n, &one, (const C *)A, lda, sA, (C *)B, ldb,
                                         sB, bc);
  }
  static rocblas_status gemm_sub(rocblas_handle h, int m, int n, int k, const double2 *A, int lda, rocblas_stride sA,
                                 const double2 *B, int ldb, rocblas_stride sB, double2 *Cm, int ldc,
                                 rocblas_stride sC, int bc) {
    const C mone(-1.0, 0.0), one(1.0, 0.0);
    return rocblas_zgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, m, n, k, &mone,
                                         (const C *)A, lda, sA, (const C *)B, ldb, sB, &one, (C *)Cm, ldc, sC, bc);
  }
  static rocblas_status trtri(rocblas_handle h, rocblas_fill uplo, rocblas_diagonal diag, int n, double2 *A, int lda,
                              rocblas_stride sA, int *info, int bc) {
    return rocsolver_ztrtri_strided_batched(h, uplo, diag, n, (C *)A, lda, sA, info, bc);
  }
};
template <>
struct Lib<float2> {
  using C = rocblas_float_complex;
  static rocblas_status getrf(rocblas_handle h, int m, int n, float2 *A, int lda, rocblas_stride sA, int *ip,
                              rocblas_stride sP, int *info, int bc) {
    return rocsolver_cgetrf_strided_batched(h, m, n, (C *)A, lda, sA, ip, sP, info, bc);
  }
  static rocblas_status trsm_ru(rocblas_handle h, int m, int n, const float2 *A, int lda, rocblas_stride sA, float2 *B,
                                int ldb, rocblas_stride sB, int bc) {
    const C one(1.0f, 0.0f);
    return rocblas_ctrsm_strided_batched(h, rocblas_side_right, rocblas_fill_upper, rocblas_operation_none,
                                         rocblas_diagonal_non_unit, m, n, &one, (const C *)A, lda, sA, (C *)B, ldb,
                                         sB, bc);
  }
  static rocblas_status gemm_sub(rocblas_handle h, int m, int n, int k, const float2 *A, int lda, rocblas_stride sA,
                                 const float2 *B, int ldb, rocblas_stride sB, float2 *Cm, int ldc, rocblas_stride sC,
                                 int bc) {
    const C mone(-1.0f, 0.0f), one(1.0f, 0.0f);
    return rocblas_cgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, m, n, k, &mone,
                                         (const C *)A, lda, sA, (const C *)B, ldb, sB, &one, (C *)Cm, ldc, sC, bc);
  }
  static rocblas_status trtri(rocblas_handle h, rocblas_fill uplo, rocblas_diagonal diag, int n, float2 *A, int lda,
                              rocblas_stride sA, int *info, int bc) {
    return rocsolver_ctrtri_strided_batched(h, uplo, diag, n, (C *)A, lda, sA, info, bc);
  }
};

}  // namespace

// ---------------------------------------------------------------- host side
struct LuFactor::Impl {
  hipStream_t st = nullptr;
  rocblas_handle blas = nullptr;
  // tree on the device
  DevBuf dbeg, uoff, foff, fvoff, uidx, umap, parent, kid_off, kids, perm, asm_pos;
  int64_t fstride = 0, fvtot = 0;
  // schedule
  std::vector<std::vector<int32_t>> small_nodes, big_nodes, level_nodes;  // per level (root excluded)
  DevBuf d_small, d_levels;                                             // node lists, concatenated
  std::vector<int64_t> small_off, level_off;
  std::vector<std::vector<int2>> ea_tasks;  // per (level * max_children + slot)
  DevBuf d_ea;
  std::vector<int64_t> ea_off;
  int ea_cb = 16;
  DevBuf d_bigoff, d_biglen;  // per big node: dbeg, nd (for the pivot conversion)
  std::vector<int32_t> big_all;
  // root
  int root = -1, nexc = 1, no_max = 0;
  std::vector<int32_t> no;
  DevBuf d_O, d_no, d_root_off, d_root_len;
  // numeric storage
  DevBuf fronts, pp, ipiv, info, zp, E, rpp, ripiv;
  // apply storage and per-ty_n task lists
  DevBuf b, y, t, x, fv, yr;
  struct ApplyTasks {
    DevBuf d;
    std::vector<int64_t> off[6];  // per kind: offsets per level (nlevels + 1)
    std::vector<int2> all;
  };
  std::map<int, ApplyTasks> atasks;  // keyed by ty_n
  ~Impl() {
    if (blas) rocblas_destroy_handle(blas);
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
    D.fstride = fstride;
    D.fvstride = fvtot * nrhs;
    return D;
  }
};

LuFactor::LuFactor() : d_(new Impl) {}
LuFactor::~LuFactor() { delete d_; }

size_t LuFactor::device_bytes_per_instance() const {
  if (!T_) return 0;
  return (size_t)d_->fstride * (prec_ == 0 ? 16 : 8);
}

void LuFactor::setup(const LuTree &T, int nexc, const std::vector<int32_t> &root_group, hipStream_t st) {
  Impl &I = *d_;
  T_ = &T;
  I.st = st;
  if (!I.blas) ROCBLAS_CHK(rocblas_create_handle(&I.blas));
  ROCBLAS_CHK(rocblas_set_stream(I.blas, st));
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
  I.fstride = T.foff[nn];
  // root
  I.root = T.last_node;
  I.nexc = std::max(1, nexc);
  if (I.root >= 0) {
    const int64_t ndr = T.nd(I.root);
    if ((int64_t)root_group.size() != ndr) throw std::invalid_argument("lu: root_group must cover the port rows");
    std::vector<std::vector<int32_t>> O(I.nexc);
    for (int64_t r = 0; r < ndr; r++) {
      if (root_group[r] < 0 || root_group[r] >= I.nexc) throw std::invalid_argument("lu: bad root group");
      for (int e = 0; e < I.nexc; e++)
        if (root_group[r] != e) O[e].push_back((int32_t)r);
    }
    I.no.assign(I.nexc, 0);
    I.no_max = 1;
    for (int e = 0; e < I.nexc; e++) {
      I.no[e] = (int32_t)O[e].size();
      I.no_max = std::max(I.no_max, I.no[e]);
    }
    std::vector<int32_t> Of((size_t)I.nexc * I.no_max, 0);
    for (int e = 0; e < I.nexc; e++) std::copy(O[e].begin(), O[e].end(), Of.begin() + (size_t)e * I.no_max);
    I.d_O.upload(Of, st);
    I.d_no.upload(I.no, st);
    std::vector<int64_t> roff(I.nexc);
    for (int e = 0; e < I.nexc; e++) roff[e] = (int64_t)e * I.no_max;
    I.d_root_off.upload(roff, st);
    I.d_root_len.upload(I.no, st);
  }
  // schedule per level
  const int L = T.nlevels;
  I.small_nodes.assign(L, {});
  I.big_nodes.assign(L, {});
  I.level_nodes.assign(L, {});
  for (int t = 0; t < nn; t++) {
    I.level_nodes[T.level[t]].push_back(t);
    if (t == I.root) continue;
    if (T.nf(t) <= kSmallFront)
      I.small_nodes[T.level[t]].push_back(t);
    else
      I.big_nodes[T.level[t]].push_back(t);
  }
  std::vector<int32_t> cat;
  I.small_off.assign(L + 1, 0);
  for (int l = 0; l < L; l++) {
    cat.insert(cat.end(), I.small_nodes[l].begin(), I.small_nodes[l].end());
    I.small_off[l + 1] = (int64_t)cat.size();
  }
  I.d_small.upload(cat, st);
  cat.clear();
  I.level_off.assign(L + 1, 0);
  for (int l = 0; l < L; l++) {
    cat.insert(cat.end(), I.level_nodes[l].begin(), I.level_nodes[l].end());
    I.level_off[l + 1] = (int64_t)cat.size();
  }
  I.d_levels.upload(cat, st);
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
  // big nodes: pivot conversion lists
  I.big_all.clear();
  for (int l = 0; l < L; l++) I.big_all.insert(I.big_all.end(), I.big_nodes[l].begin(), I.big_nodes[l].end());
  I.atasks.clear();
  HIPCHK(hipStreamSynchronize(st));
}

template <typename T>
static void factor_impl(LuFactor::Impl &I, const LuTree &Tr, int ninst, const double *vals, int64_t vstride,
                        LuFactorStats *stats) {
  hipStream_t st = I.st;
  const int nn = Tr.nnodes;
  const int64_t fs = I.fstride;
  I.fronts.alloc((size_t)ninst * fs * sizeof(T));
  I.pp.alloc((size_t)ninst * std::max<int64_t>(Tr.n, 1) * 4);
  I.ipiv.alloc((size_t)ninst * std::max<int64_t>(Tr.n, 1) * 4);
  const int nbig = (int)I.big_all.size();
  I.info.alloc((size_t)ninst * (2 * nbig + 3 * I.nexc + 1) * 4);
  I.zp.alloc(8);
  HIPCHK(hipMemsetAsync(I.info.p, 0, (size_t)ninst * (2 * nbig + 3 * I.nexc + 1) * 4, st));
  HIPCHK(hipMemsetAsync(I.zp.p, 0, 8, st));
  T *F = I.fronts.as<T>();
  LuDev D = I.dev(Tr, 1);
  hipLaunchKernelGGL((k_lu_zero<T>), dim3(grid_for((int64_t)ninst * fs, 65536)), dim3(256), 0, st,
                     (int64_t)ninst * fs, F);
  hipLaunchKernelGGL((k_lu_assemble<T>), dim3(grid_for((int64_t)Tr.asm_pos.size(), 16384), ninst), dim3(256), 0, st,
                     (int64_t)Tr.asm_pos.size(), I.asm_pos.as<int64_t>(), (const double2 *)vals, vstride, F, fs);
  LU_LAUNCH_CHECK();
  const int MC = std::max(1, Tr.max_children);
  int info_k = 0;
  int *info = I.info.as<int>();
  for (int l = 0; l < Tr.nlevels; l++) {
    for (int s = 0; s < MC; s++) {
      const int64_t a = I.ea_off[(size_t)l * MC + s], b = I.ea_off[(size_t)l * MC + s + 1];
      if (b > a)
        hipLaunchKernelGGL((k_lu_extend_add<T>), dim3((unsigned)(b - a), ninst), dim3(256), 0, st,
                           I.d_ea.as<int2>() + a, D, F, I.ea_cb);
    }
    LU_LAUNCH_CHECK();
    const int64_t s0 = I.small_off[l], s1 = I.small_off[l + 1];
    if (s1 > s0)
      hipLaunchKernelGGL((k_lu_front_small<T>), dim3((unsigned)(s1 - s0), ninst), dim3(256), 0, st,
                         I.d_small.as<int32_t>() + s0, D, F, I.pp.as<int32_t>(),
                         I.zp.as<unsigned long long>());
    LU_LAUNCH_CHECK();
    for (int t : I.big_nodes[l]) {
      const int nd = (int)Tr.nd(t), nu = (int)Tr.nu(t), nf = nd + nu;
      T *Ft = F + Tr.foff[t];
      int *ip = I.ipiv.as<int>() + Tr.dbeg[t];
      ROCBLAS_CHK(Lib<T>::getrf(I.blas, nd, nf, Ft, nf, fs, ip, Tr.n, info + (size_t)info_k * ninst, ninst));
      info_k++;
      if (nu > 0) {
        ROCBLAS_CHK(Lib<T>::trsm_ru(I.blas, nu, nd, Ft, nf, fs, Ft + nd, nf, fs, ninst));
        ROCBLAS_CHK(Lib<T>::gemm_sub(I.blas, nu, nu, nd, Ft + nd, nf, fs, Ft + (int64_t)nf * nd, nf, fs,
                                     Ft + nd + (int64_t)nf * nd, nf, fs, ninst));
      }
      ROCBLAS_CHK(Lib<T>::trtri(I.blas, rocblas_fill_upper, rocblas_diagonal_non_unit, nd, Ft, nf, fs,
                                info + (size_t)info_k * ninst, ninst));
      info_k++;
      ROCBLAS_CHK(Lib<T>::trtri(I.blas, rocblas_fill_lower, rocblas_diagonal_unit, nd, Ft, nf, fs,
                                info + (size_t)info_k * ninst, ninst));
      // the unit-lower trtri reports no singularity; its slot is reused
    }
    if (!I.big_nodes[l].empty()) {
      std::vector<int64_t> off;
      std::vector<int32_t> len;
      int mx = 0;
      for (int t : I.big_nodes[l]) {
        off.push_back(Tr.dbeg[t]);
        len.push_back((int32_t)Tr.nd(t));
        mx = std::max(mx, (int)Tr.nd(t));
      }
      DevBuf doff, dlen;
      doff.upload(off, st);
      dlen.upload(len, st);
      hipLaunchKernelGGL(k_lu_ipiv_to_perm, dim3((unsigned)off.size(), ninst), dim3(256), (size_t)mx * 4, st,
                         doff.as<int64_t>(), dlen.as<int32_t>(), I.ipiv.as<int32_t>(), I.pp.as<int32_t>(), Tr.n);
      LU_LAUNCH_CHECK();
      HIPCHK(hipStreamSynchronize(st));  // doff/dlen go out of scope
    }
  }
  if (I.root >= 0) {
    const int r = I.root;
    const int nfr = (int)Tr.nd(r);
    const int64_t es = (int64_t)I.no_max * I.no_max;
    I.E.alloc((size_t)ninst * I.nexc * es * sizeof(T));
    I.rpp.alloc((size_t)ninst * I.nexc * I.no_max * 4);
    I.ripiv.alloc((size_t)ninst * I.nexc * I.no_max * 4);
    T *E = I.E.as<T>();
    for (int e = 0; e < I.nexc; e++) {
      const int no = I.no[e];
      if (no == 0) continue;
      T *Ee = E + (int64_t)e * es;
      hipLaunchKernelGGL((k_lu_root_gather<T>), dim3(grid_for((int64_t)no * no, 16384), ninst), dim3(256), 0, st,
                         F + Tr.foff[r], fs, nfr, I.d_O.as<int32_t>() + (int64_t)e * I.no_max, no, Ee,
                         (int64_t)I.nexc * es);
      LU_LAUNCH_CHECK();
      int *rip = I.ripiv.as<int>() + (int64_t)e * I.no_max;
      ROCBLAS_CHK(Lib<T>::getrf(I.blas, no, no, Ee, no, (int64_t)I.nexc * es, rip, (int64_t)I.nexc * I.no_max,
                                info + (size_t)info_k * ninst, ninst));
      info_k++;
      ROCBLAS_CHK(Lib<T>::trtri(I.blas, rocblas_fill_upper, rocblas_diagonal_non_unit, no, Ee, no,
                                (int64_t)I.nexc * es, info + (size_t)info_k * ninst, ninst));
      info_k++;
      ROCBLAS_CHK(Lib<T>::trtri(I.blas, rocblas_fill_lower, rocblas_diagonal_unit, no, Ee, no, (int64_t)I.nexc * es,
                                info + (size_t)info_k * ninst, ninst));
    }
    // pivot permutations of all excitation blocks (matrix stride nexc * no_max)
    hipLaunchKernelGGL(k_lu_ipiv_to_perm, dim3(I.nexc, ninst), dim3(256), (size_t)I.no_max * 4, st,
                       I.d_root_off.as<int64_t>(), I.d_root_len.as<int32_t>(), I.ripiv.as<int32_t>(),
                       I.rpp.as<int32_t>(), (int64_t)I.nexc * I.no_max);
    LU_LAUNCH_CHECK();
  }
  std::vector<int> hinfo((size_t)info_k * ninst);
  if (!hinfo.empty()) HIPCHK(hipMemcpyAsync(hinfo.data(), info, hinfo.size() * 4, hipMemcpyDeviceToHost, st));
  unsigned long long zp = 0;
  HIPCHK(hipMemcpyAsync(&zp, I.zp.p, 8, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (stats) {
    stats->zero_pivots = (int64_t)zp;
    stats->lib_info = 0;
    for (int v : hinfo) stats->lib_info += v != 0;
  }
  (void)nn;
}

void LuFactor::factor(int prec, int ninst, const double *vals, int64_t vals_stride, LuFactorStats *stats) {
  if (!T_) throw std::logic_error("lu: setup first");
  if (ninst < 1) throw std::invalid_argument("lu: ninst must be >= 1");
  const auto t0 = std::chrono::steady_clock::now();
  prec_ = prec;
  ninst_ = ninst;
  if (prec == 0)
    factor_impl<double2>(*d_, *T_, ninst, vals, vals_stride, stats);
  else
    factor_impl<float2>(*d_, *T_, ninst, vals, vals_stride, stats);
  if (stats) stats->seconds_factor = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

template <typename T, typename X>
static void solve_impl(LuFactor::Impl &I, const LuTree &Tr, int ninst, int nrhs, const X *v, int64_t ldv, X *z) {
  if (nrhs > 4 * kMaxRhsThread) throw std::invalid_argument("lu: at most 64 right-hand sides per matrix");
  hipStream_t st = I.st;
  const int64_t n = Tr.n, S = (int64_t)ninst * nrhs;
  const int ty_n = (nrhs + kMaxRhsThread - 1) / kMaxRhsThread;
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
    LuFactor::Impl::ApplyTasks A;
    const int L = Tr.nlevels;
    for (int kind = 0; kind < 6; kind++) A.off[kind].assign(L + 1, 0);
    // layout: kind-major, then level
    for (int kind = 0; kind < 6; kind++) {
      for (int l = 0; l < L; l++) {
        A.off[kind][l] = (int64_t)A.all.size();
        if (kind <= K_BWD_U) {
          for (int t : I.level_nodes[l]) {
            if (t == I.root) continue;
            const int64_t rows = kind == K_FWD_L21 ? Tr.nu(t) : Tr.nd(t);
            for (int64_t r0 = 0; r0 < rows; r0 += rows_wg) A.all.push_back(make_int2(t, (int)r0));
          }
        } else if (I.root >= 0 && Tr.level[I.root] == l) {
          for (int e = 0; e < I.nexc; e++)
            for (int64_t r0 = 0; r0 < I.no[e]; r0 += rows_wg) A.all.push_back(make_int2(e, (int)r0));
        }
      }
      A.off[kind][L] = (int64_t)A.all.size();
    }
    it = I.atasks.emplace(std::piecewise_construct, std::forward_as_tuple(ty_n), std::forward_as_tuple()).first;
    it->second.all = std::move(A.all);
    for (int kind = 0; kind < 6; kind++) it->second.off[kind] = std::move(A.off[kind]);
    it->second.d.upload(it->second.all, st);
  }
  auto &A = it->second;
  LuDev D = I.dev(Tr, nrhs);
  RootDev Rt{};
  Rt.node = I.root;
  Rt.nexc = I.nexc;
  Rt.O = I.d_O.as<int32_t>();
  Rt.no = I.d_no.as<int32_t>();
  Rt.no_max = std::max(1, I.no_max);
  Rt.estride = (int64_t)Rt.no_max * Rt.no_max;
  Rt.rpp = I.rpp.as<int32_t>();
  const T *F = I.fronts.as<T>();
  const T *E = I.E.as<T>();
  const int32_t *pp = I.pp.as<int32_t>();
  double2 *b = I.b.as<double2>(), *y = I.y.as<double2>(), *tt = I.t.as<double2>(), *x = I.x.as<double2>(),
          *fv = I.fv.as<double2>(), *yr = I.yr.as<double2>();
  hipLaunchKernelGGL((k_lu_gather_rhs<X>), dim3(grid_for(n * nrhs), ninst), dim3(256), 0, st, n, nrhs,
                     I.perm.as<int32_t>(), v, ldv, b);
  LU_LAUNCH_CHECK();
  auto launch = [&](int kind, int l) {
    const int64_t a0 = A.off[kind][l], a1 = A.off[kind][l + 1];
    if (a1 <= a0) return;
    const int2 *tk = A.d.as<int2>() + a0;
    const dim3 g((unsigned)(a1 - a0), ninst);
    switch (kind) {
      case K_FWD_L:
        hipLaunchKernelGGL((k_lu_apply<T, K_FWD_L>), g, dim3(256), 0, st, tk, D, Rt, nrhs, ty_n, F, E, pp, fv, y, tt,
                           x, yr);
        break;
      case K_FWD_L21:
        hipLaunchKernelGGL((k_lu_apply<T, K_FWD_L21>), g, dim3(256), 0, st, tk, D, Rt, nrhs, ty_n, F, E, pp, fv, y,
                           tt, x, yr);
        break;
      case K_BWD_U12:
        hipLaunchKernelGGL((k_lu_apply<T, K_BWD_U12>), g, dim3(256), 0, st, tk, D, Rt, nrhs, ty_n, F, E, pp, fv, y,
                           tt, x, yr);
        break;
      case K_BWD_U:
        hipLaunchKernelGGL((k_lu_apply<T, K_BWD_U>), g, dim3(256), 0, st, tk, D, Rt, nrhs, ty_n, F, E, pp, fv, y, tt,
                           x, yr);
        break;
      case K_ROOT_L:
        hipLaunchKernelGGL((k_lu_apply<T, K_ROOT_L>), g, dim3(256), 0, st, tk, D, Rt, nrhs, ty_n, F, E, pp, fv, y,
                           tt, x, yr);
        break;
      default:
        hipLaunchKernelGGL((k_lu_apply<T, K_ROOT_U>), g, dim3(256), 0, st, tk, D, Rt, nrhs, ty_n, F, E, pp, fv, y,
                           tt, x, yr);
        break;
    }
    LU_LAUNCH_CHECK();
  };
  const int L = Tr.nlevels;
  for (int l = 0; l < L; l++) {
    const int64_t a0 = I.level_off[l], a1 = I.level_off[l + 1];
    if (a1 > a0)
      hipLaunchKernelGGL(k_lu_fwd_gather, dim3((unsigned)(a1 - a0), ninst), dim3(256), 0, st,
                         I.d_levels.as<int32_t>() + a0, D, nrhs, b, fv);
    LU_LAUNCH_CHECK();
    launch(K_FWD_L, l);
    launch(K_FWD_L21, l);
    launch(K_ROOT_L, l);
  }
  if (I.root >= 0) {
    hipLaunchKernelGGL(k_lu_zero_rows, dim3(grid_for(Tr.nd(I.root) * nrhs), ninst), dim3(256), 0, st,
                       (int64_t)Tr.dbeg[I.root], (int64_t)Tr.nd(I.root), nrhs, n * nrhs, x);
    LU_LAUNCH_CHECK();
  }
  for (int l = L - 1; l >= 0; l--) {
    launch(K_ROOT_U, l);
    launch(K_BWD_U12, l);
    launch(K_BWD_U, l);
  }
  const int64_t ndof = Tr.ndof;
  hipLaunchKernelGGL((k_lu_zero<X>), dim3(grid_for(ndof * S, 65536)), dim3(256), 0, st, ndof * S, z);
  hipLaunchKernelGGL((k_lu_scatter<X>), dim3(grid_for(n * nrhs), ninst), dim3(256), 0, st, n, nrhs, S,
                     I.perm.as<int32_t>(), x, z);
  LU_LAUNCH_CHECK();
}

template <typename X>
void LuFactor::solve(int nrhs, const X *v, int64_t ldv, X *z) {
  if (!T_ || ninst_ < 1) throw std::logic_error("lu: factor first");
  if (prec_ == 0)
    solve_impl<double2, X>(*d_, *T_, ninst_, nrhs, v, ldv, z);
  else
    solve_impl<float2, X>(*d_, *T_, ninst_, nrhs, v, ldv, z);
}
template void LuFactor::solve<double2>(int, const double2 *, int64_t, double2 *);
template void LuFactor::solve<float2>(int, const float2 *, int64_t, float2 *);

}  // namespace opem
