// Device kernels of the per-frequency solve path for gfx950 (wave64, CDNA4).
// See kernels.h for the contracts.
#include "kernels.h"

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <type_traits>

namespace opem {

#define HIP_LAUNCH_CHECK()                                                                         \
  do {                                                                                             \
    hipError_t e_ = hipGetLastError();                                                             \
    if (e_ != hipSuccess) throw std::runtime_error(std::string("kernel launch: ") + hipGetErrorString(e_)); \
  } while (0)

// ------------------------------------------------------------------ complex helpers
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cconjmul(double2 a, double2 b) {  // conj(a)*b
  return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
template <typename T>
__device__ __forceinline__ double2 to_d2(T v) {
  return make_double2((double)v.x, (double)v.y);
}
template <typename T>
__device__ __forceinline__ T from_d2(double2 v);
template <>
__device__ __forceinline__ double2 from_d2<double2>(double2 v) {
  return v;
}
template <>
__device__ __forceinline__ float2 from_d2<float2>(double2 v) {
  return make_float2((float)v.x, (float)v.y);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ double wave_sum_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// Blocks are dispatched round-robin over the 8 XCDs (block b -> XCD b % 8).  Remapping the linear block
// id so that each XCD receives one contiguous range of logical blocks keeps the x-vector slice a range of
// rows touches in that XCD's own L2.
__device__ __forceinline__ int64_t xcd_logical(int64_t b, int64_t nb) {
  const int64_t q = nb / 8, rem = nb % 8, k = b % 8, i = b / 8;
  return k * q + (k < rem ? k : rem) + i;
}

constexpr int kRedThreads = 256;
constexpr int kRedElems = 4;
constexpr int kChunk = kRedThreads * kRedElems;

int reduce_blocks(int64_t n) {
  int64_t b = (n + kChunk - 1) / kChunk;
  return (int)(b < 1 ? 1 : b);
}

// ------------------------------------------------------------------ volume assembly
constexpr int kMaxSlot = 8;

__global__ __launch_bounds__(256) void k_assemble_volume(
    int64_t ndof, const int32_t *__restrict__ rowptr, const int32_t *__restrict__ colind,
    const int32_t *__restrict__ d2e_ptr, const int32_t *__restrict__ d2e, const int32_t *__restrict__ elem_slot,
    const double *__restrict__ gm, const double *__restrict__ gk, const double *__restrict__ sdet,
    const double *__restrict__ TM, const double *__restrict__ TK, const double *__restrict__ TQ, int nd,
    int nslot, int64_t nnz, double *__restrict__ K, double *__restrict__ M, double *__restrict__ Q) {
  const int lane = threadIdx.x & 63;
  const int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (r >= ndof) return;
  const int rb = d2e_ptr[r], re = d2e_ptr[r + 1];
  const int ndd = nd * nd;
  for (int nz = rowptr[r] + lane; nz < rowptr[r + 1]; nz += 64) {
    const int c = colind[nz];
    const int cb = d2e_ptr[c], ce = d2e_ptr[c + 1];
    double k[kMaxSlot], m[kMaxSlot], q[kMaxSlot];
#pragma unroll
    for (int t = 0; t < kMaxSlot; t++) k[t] = m[t] = q[t] = 0.0;
    int i = rb, j = cb;
    while (i < re && j < ce) {
      const int pr = d2e[i], pc = d2e[j];
      const int er = pr / nd, ec = pc / nd;
      if (er < ec) {
        i++;
      } else if (ec < er) {
        j++;
      } else {
        const int a = pr - er * nd, b = pc - ec * nd;
        const int sl = elem_slot[er];
        double kk = 0.0, mm = 0.0;
#pragma unroll
        for (int t = 0; t < 6; t++) {
          kk += gk[(int64_t)er * 6 + t] * TK[t * ndd + a * nd + b];
          mm += gm[(int64_t)er * 6 + t] * TM[t * ndd + a * nd + b];
        }
        const double qq = sdet[er] * TQ[a * nd + b];
#pragma unroll
        for (int t = 0; t < kMaxSlot; t++)
          if (t == sl) {
            k[t] += kk;
            m[t] += mm;
            q[t] += qq;
          }
        i++;
        j++;
      }
    }
    for (int t = 0; t < nslot; t++) {
      K[t * nnz + nz] = k[t];
      M[t * nnz + nz] = m[t];
      Q[t * nnz + nz] = q[t];
    }
  }
}

void launch_assemble_volume(hipStream_t st, int64_t ndof, const int32_t *rowptr, const int32_t *colind,
                            const int32_t *d2e_ptr, const int32_t *d2e, const int32_t *elem_slot,
                            const double *gm, const double *gk, const double *sdet, const double *TM,
                            const double *TK, const double *TQ, int nd, int nslot, int64_t nnz, double *K,
                            double *M, double *Q) {
  if (nslot > kMaxSlot) throw std::runtime_error("too many volume attributes (max 8)");
  dim3 grid((unsigned)((ndof + 3) / 4));
  hipLaunchKernelGGL(k_assemble_volume, grid, dim3(256), 0, st, ndof, rowptr, colind, d2e_ptr, d2e, elem_slot,
                     gm, gk, sdet, TM, TK, TQ, nd, nslot, nnz, K, M, Q);
  HIP_LAUNCH_CHECK();
}

// ------------------------------------------------------------------ per-frequency values
__global__ __launch_bounds__(256) void k_combine(int nf, int64_t nnz, int nslot, const double *__restrict__ cK,
                                                 const double *__restrict__ cM, const double *__restrict__ K,
                                                 const double *__restrict__ M, double2 *__restrict__ a64,
                                                 float2 *__restrict__ a32) {
  for (int64_t nz = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; nz < nnz;
       nz += (int64_t)gridDim.x * blockDim.x) {
    double kv[kMaxSlot], mv[kMaxSlot];
    for (int t = 0; t < nslot; t++) {
      kv[t] = K[t * nnz + nz];
      mv[t] = M[t * nnz + nz];
    }
    for (int f = 0; f < nf; f++) {
      double re = 0.0, im = 0.0;
      for (int t = 0; t < nslot; t++) {
        const double *ck = cK + ((int64_t)f * nslot + t) * 2;
        const double *cm = cM + ((int64_t)f * nslot + t) * 2;
        re += ck[0] * kv[t] + cm[0] * mv[t];
        im += ck[1] * kv[t] + cm[1] * mv[t];
      }
      a64[(int64_t)f * nnz + nz] = make_double2(re, im);
      if (a32) a32[(int64_t)f * nnz + nz] = make_float2((float)re, (float)im);
    }
  }
}

__global__ void k_overlay(int nf, int64_t nnz, int64_t nov, int nterm, const int32_t *__restrict__ ov_nz,
                          const double *__restrict__ ov_val, const double *__restrict__ cT, double2 *a64,
                          float2 *a32) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int f = blockIdx.y;
  if (i >= nov) return;
  double re = 0.0, im = 0.0;
  for (int t = 0; t < nterm; t++) {
    const double v = ov_val[(int64_t)t * nov + i];
    re += cT[((int64_t)f * nterm + t) * 2] * v;
    im += cT[((int64_t)f * nterm + t) * 2 + 1] * v;
  }
  const int64_t nz = (int64_t)f * nnz + ov_nz[i];
  double2 a = a64[nz];
  a.x += re;
  a.y += im;
  a64[nz] = a;
  if (a32) a32[nz] = make_float2((float)a.x, (float)a.y);
}

void launch_combine(hipStream_t st, int nf, int64_t nnz, int nslot, const double *cK, const double *cM,
                    const double *K, const double *M, int64_t nov, int nterm, const int32_t *ov_nz,
                    const double *ov_val, const double *cT, double *a64, float *a32) {
  int64_t blocks = (nnz + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_combine, dim3((unsigned)blocks), dim3(256), 0, st, nf, nnz, nslot, cK, cM, K, M,
                     (double2 *)a64, (float2 *)a32);
  HIP_LAUNCH_CHECK();
  if (nov > 0 && nterm > 0) {
    hipLaunchKernelGGL(k_overlay, dim3((unsigned)((nov + 255) / 256), nf), dim3(256), 0, st, nf, nnz, nov, nterm,
                       ov_nz, ov_val, cT, (double2 *)a64, (float2 *)a32);
    HIP_LAUNCH_CHECK();
  }
}

// ------------------------------------------------------------------ SpMV
// LPR lanes cooperate on one row; R right-hand sides share each value load.
template <typename V, typename X, int LPR, int R>
__global__ __launch_bounds__(256) void k_spmv(int64_t n, const int32_t *__restrict__ rowptr,
                                              const int32_t *__restrict__ colind, const V *__restrict__ vals,
                                              int64_t nnz, const uint8_t *__restrict__ rowmask,
                                              const X *__restrict__ x, X *__restrict__ y, const X *__restrict__ b,
                                              const int32_t *__restrict__ active, int mode, int64_t vstride,
                                              int xcd_remap) {
  int64_t bx = blockIdx.x, by = blockIdx.y;
  if (xcd_remap) {
    const int64_t nbx = gridDim.x;
    const int64_t lg = xcd_logical(by * nbx + bx, nbx * gridDim.y);
    bx = lg % nbx;
    by = lg / nbx;
  }
  const int g = (int)by;
  bool any = false;
#pragma unroll
  for (int r = 0; r < R; r++) any |= (active == nullptr) || active[g * R + r];
  if (!any) return;
  const int sub = threadIdx.x % LPR;
  const int64_t row = (bx * blockDim.x + threadIdx.x) / LPR;
  if (row >= n) return;
  const V *__restrict__ A = vals + (int64_t)g * vstride;
  const int beg = rowptr[row], end = rowptr[row + 1];
  double2 acc[R];
#pragma unroll
  for (int r = 0; r < R; r++) acc[r] = make_double2(0.0, 0.0);
  for (int k = beg + sub; k < end; k += LPR) {
    const int c = colind[k];
    const V a = A[k];
#pragma unroll
    for (int r = 0; r < R; r++) {
      const X xv = x[((int64_t)g * R + r) * n + c];
      acc[r].x += (double)a.x * (double)xv.x - (double)a.y * (double)xv.y;
      acc[r].y += (double)a.x * (double)xv.y + (double)a.y * (double)xv.x;
    }
  }
#pragma unroll
  for (int r = 0; r < R; r++) {
#pragma unroll
    for (int o = LPR / 2; o > 0; o >>= 1) {
      acc[r].x += __shfl_xor(acc[r].x, o, LPR);
      acc[r].y += __shfl_xor(acc[r].y, o, LPR);
    }
  }
  if (sub == 0) {
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int64_t s = (int64_t)g * R + r;
      if (active && !active[s]) continue;
      const bool free_row = rowmask[(int64_t)r * n + row] != 0;
      double2 out = make_double2(0.0, 0.0);
      if (free_row) {
        if (mode == 0) {
          out = acc[r];
        } else {
          double2 bv = b ? to_d2(b[s * n + row]) : make_double2(0.0, 0.0);
          out = make_double2(bv.x - acc[r].x, bv.y - acc[r].y);
        }
      }
      y[s * n + row] = from_d2<X>(out);
    }
  }
}

int spmv_pair_enabled() {
  static int v = -1;
  if (v < 0) {
    const char *e = getenv("OPEM_SPMV_PAIR");
    v = e ? atoi(e) : 1;
  }
  return v;
}

int spmv_wide_enabled() {
  static int v = -1;
  if (v < 0) {
    const char *e = getenv("OPEM_SPMV_WIDE");
    v = e ? atoi(e) : 0;
  }
  return v;
}

float patch_split_kappa() {
  static float v = -1.f;
  if (v < 0.f) {
    const char *e = getenv("OPEM_PATCH_SPLIT_KAPPA");
    v = e ? (float)atof(e) : 20.f;
    if (!(v >= 0.f)) v = 20.f;
  }
  return v;
}

int patch_mfma_enabled() {
  static int v = -1;
  if (v < 0) {
    const char *e = getenv("OPEM_PATCH_MFMA");
    v = e ? atoi(e) : 1;
  }
  return v;
}

int xcd_remap_enabled() {
  static int v = -1;
  if (v < 0) {
    const char *e = getenv("OPEM_XCD_REMAP");
    v = e ? atoi(e) : 0;
  }
  return v;
}

template <typename V, typename X, int LPR>
static void spmv_r(hipStream_t st, int ngroups, int R, int64_t n, const int32_t *rowptr, const int32_t *colind,
                   const V *vals, int64_t nnz, const uint8_t *rowmask, const X *x, X *y, const X *b,
                   const int32_t *active, int mode, int64_t vstride) {
  dim3 grid((unsigned)((n * LPR + 255) / 256), ngroups);
  switch (R) {
    case 1:
      hipLaunchKernelGGL((k_spmv<V, X, LPR, 1>), grid, dim3(256), 0, st, n, rowptr, colind, vals, nnz, rowmask, x,
                         y, b, active, mode, vstride, xcd_remap_enabled());
      break;
    case 2:
      hipLaunchKernelGGL((k_spmv<V, X, LPR, 2>), grid, dim3(256), 0, st, n, rowptr, colind, vals, nnz, rowmask, x,
                         y, b, active, mode, vstride, xcd_remap_enabled());
      break;
    case 3:
      hipLaunchKernelGGL((k_spmv<V, X, LPR, 3>), grid, dim3(256), 0, st, n, rowptr, colind, vals, nnz, rowmask, x,
                         y, b, active, mode, vstride, xcd_remap_enabled());
      break;
    case 4:
      hipLaunchKernelGGL((k_spmv<V, X, LPR, 4>), grid, dim3(256), 0, st, n, rowptr, colind, vals, nnz, rowmask, x,
                         y, b, active, mode, vstride, xcd_remap_enabled());
      break;
    default:
      throw std::runtime_error("spmv: R must be 1..4");
  }
  HIP_LAUNCH_CHECK();
}

template <typename V, typename X>
void launch_spmv(hipStream_t st, int ngroups, int R, int64_t n, const int32_t *rowptr, const int32_t *colind,
                 const V *vals, int64_t nnz, const uint8_t *rowmask, const X *x, X *y, const X *b,
                 const int32_t *active, int mode, int lanes_per_row, int64_t vstride) {
  switch (lanes_per_row) {
    case 4:
      spmv_r<V, X, 4>(st, ngroups, R, n, rowptr, colind, vals, nnz, rowmask, x, y, b, active, mode, vstride);
      break;
    case 8:
      spmv_r<V, X, 8>(st, ngroups, R, n, rowptr, colind, vals, nnz, rowmask, x, y, b, active, mode, vstride);
      break;
    case 16:
      spmv_r<V, X, 16>(st, ngroups, R, n, rowptr, colind, vals, nnz, rowmask, x, y, b, active, mode, vstride);
      break;
    case 32:
      spmv_r<V, X, 32>(st, ngroups, R, n, rowptr, colind, vals, nnz, rowmask, x, y, b, active, mode, vstride);
      break;
    case 64:
      spmv_r<V, X, 64>(st, ngroups, R, n, rowptr, colind, vals, nnz, rowmask, x, y, b, active, mode, vstride);
      break;
    default:
      throw std::runtime_error("spmv: lanes_per_row must be 4..64");
  }
}

template void launch_spmv<double2, double2>(hipStream_t, int, int, int64_t, const int32_t *, const int32_t *,
                                            const double2 *, int64_t, const uint8_t *, const double2 *, double2 *,
                                            const double2 *, const int32_t *, int, int, int64_t);
template void launch_spmv<float2, float2>(hipStream_t, int, int, int64_t, const int32_t *, const int32_t *,
                                          const float2 *, int64_t, const uint8_t *, const float2 *, float2 *,
                                          const float2 *, const int32_t *, int, int, int64_t);

// ------------------------------------------------------------------ single-matrix complex CSR SpMV
// One explicit complex matrix, one vector.  A row group is G consecutive rows with one column list (ND_p
// numbers the DOFs of an edge, face or interior consecutively and they share their elements: G = 2 for ND2,
// 3 for ND3, checked on the host); row g of the group keeps its values at beg + g*len.  LPR lanes share a
// group, so each column index and each gathered x value serves G rows.  Per pass a lane issues U column
// loads, G*U value loads and then U gathers of x before any arithmetic; positions past the row end are
// clamped to its last entry and their products dropped by a select, so no load is conditional.
template <typename V, int G, int LPR, int U>
__global__ __launch_bounds__(256) void k_csr_spmv(int64_t n, const int32_t *__restrict__ rowptr,
                                                  const int32_t *__restrict__ colind, const V *__restrict__ vals,
                                                  const uint8_t *__restrict__ rowmask, const V *__restrict__ x,
                                                  V *__restrict__ y, int xcd_remap) {
  const int sub = threadIdx.x % LPR;
  const int64_t bid = xcd_remap ? xcd_logical(blockIdx.x, gridDim.x) : (int64_t)blockIdx.x;
  const int64_t r0 = (bid * blockDim.x + threadIdx.x) / LPR * G;
  if (r0 >= n) return;  // 256 % LPR == 0: a row group leaves together
  const int beg = rowptr[r0], len = rowptr[r0 + 1] - beg;
  double re[G], im[G];
#pragma unroll
  for (int g = 0; g < G; g++) re[g] = im[g] = 0.0;
  for (int j0 = 0; j0 < len; j0 += LPR * U) {
    int c[U];
    V a[G][U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int j = j0 + sub + u * LPR;
      const int k = beg + (j < len ? j : len - 1);
      c[u] = colind[k];
#pragma unroll
      for (int g = 0; g < G; g++) a[g][u] = vals[k + g * len];
    }
    V xv[U];
#pragma unroll
    for (int u = 0; u < U; u++) xv[u] = x[c[u]];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const bool ok = j0 + sub + u * LPR < len;
#pragma unroll
      for (int g = 0; g < G; g++) {
        const double pr = (double)a[g][u].x * (double)xv[u].x - (double)a[g][u].y * (double)xv[u].y;
        const double pi = (double)a[g][u].x * (double)xv[u].y + (double)a[g][u].y * (double)xv[u].x;
        re[g] += ok ? pr : 0.0;
        im[g] += ok ? pi : 0.0;
      }
    }
  }
#pragma unroll
  for (int g = 0; g < G; g++) {
#pragma unroll
    for (int o = LPR / 2; o > 0; o >>= 1) {
      re[g] += __shfl_xor(re[g], o, LPR);
      im[g] += __shfl_xor(im[g], o, LPR);
    }
  }
  if (sub < G) {  // after the butterfly every lane holds every row's sum; lane g writes row r0 + g
    double rr = re[0], ii = im[0];
#pragma unroll
    for (int g = 1; g < G; g++)
      if (sub == g) {
        rr = re[g];
        ii = im[g];
      }
    const bool free_row = rowmask == nullptr || rowmask[r0 + sub] != 0;
    V out;
    out.x = (decltype(out.x))(free_row ? rr : 0.0);
    out.y = (decltype(out.y))(free_row ? ii : 0.0);
    y[r0 + sub] = out;
  }
}

template <typename V, int G, int LPR, int U>
static void csr_spmv_go(hipStream_t st, int64_t n, const int32_t *rowptr, const int32_t *colind, const V *vals,
                        const uint8_t *rowmask, const V *x, V *y) {
  const int64_t groups = n / G;
  hipLaunchKernelGGL((k_csr_spmv<V, G, LPR, U>), dim3((unsigned)((groups * LPR + 255) / 256)), dim3(256), 0, st, n,
                     rowptr, colind, vals, rowmask, x, y, xcd_remap_enabled());
}

template <typename V>
void launch_csr_spmv(hipStream_t st, int64_t n, const int32_t *rowptr, const int32_t *colind, const V *vals,
                     const uint8_t *rowmask, const V *x, V *y, int row_group, int lanes_per_row) {
  if (n <= 0) return;
  if (n > (int64_t)(1u << 31) / 64) throw std::runtime_error("csr spmv: too many rows for one launch");
  if (row_group < 1 || row_group > 3 || n % row_group)
    throw std::runtime_error("csr spmv: row group must be 1, 2 or 3 and divide the row count");
  const int key = row_group * 100 + lanes_per_row;
  switch (key) {
    case 104: csr_spmv_go<V, 1, 4, 4>(st, n, rowptr, colind, vals, rowmask, x, y); break;
    case 108: csr_spmv_go<V, 1, 8, 4>(st, n, rowptr, colind, vals, rowmask, x, y); break;
    case 116: csr_spmv_go<V, 1, 16, 4>(st, n, rowptr, colind, vals, rowmask, x, y); break;
    case 132: csr_spmv_go<V, 1, 32, 2>(st, n, rowptr, colind, vals, rowmask, x, y); break;
    case 164: csr_spmv_go<V, 1, 64, 2>(st, n, rowptr, colind, vals, rowmask, x, y); break;
    case 208: csr_spmv_go<V, 2, 8, 2>(st, n, rowptr, colind, vals, rowmask, x, y); break;
    case 216: csr_spmv_go<V, 2, 16, 2>(st, n, rowptr, colind, vals, rowmask, x, y); break;
    case 232: csr_spmv_go<V, 2, 32, 2>(st, n, rowptr, colind, vals, rowmask, x, y); break;
    case 308: csr_spmv_go<V, 3, 8, 2>(st, n, rowptr, colind, vals, rowmask, x, y); break;
    case 316: csr_spmv_go<V, 3, 16, 2>(st, n, rowptr, colind, vals, rowmask, x, y); break;
    case 332: csr_spmv_go<V, 3, 32, 2>(st, n, rowptr, colind, vals, rowmask, x, y); break;
    default:
      throw std::runtime_error("csr spmv: lanes_per_row must be 4..64 for single rows, 8, 16 or 32 for row groups");
  }
  HIP_LAUNCH_CHECK();
}
template void launch_csr_spmv<double2>(hipStream_t, int64_t, const int32_t *, const int32_t *, const double2 *,
                                       const uint8_t *, const double2 *, double2 *, int, int);
template void launch_csr_spmv<float2>(hipStream_t, int64_t, const int32_t *, const int32_t *, const float2 *,
                                      const uint8_t *, const float2 *, float2 *, int, int);

// Largest G in {3, 2} such that G divides n and every G consecutive rows starting at a multiple of G have one
// column list; 1 otherwise.
int csr_row_group(int64_t n, const int32_t *rowptr, const int32_t *colind) {
  for (int G = 3; G >= 2; G--) {
    if (n % G) continue;
    bool ok = true;
    for (int64_t r0 = 0; ok && r0 < n; r0 += G) {
      const int32_t beg = rowptr[r0], len = rowptr[r0 + 1] - beg;
      for (int g = 1; ok && g < G; g++) {
        const int32_t bg = rowptr[r0 + g];
        if (rowptr[r0 + g + 1] - bg != len || bg != beg + g * len) {
          ok = false;
          break;
        }
        for (int32_t j = 0; j < len; j++)
          if (colind[bg + j] != colind[beg + j]) {
            ok = false;
            break;
          }
      }
    }
    if (ok) return G;
  }
  return 1;
}

// ------------------------------------------------------------------ batched sweep SpMV
// y[s][row] = rowmask[s % R][row] * (+/-) sum_j (A_s)_{row j} x[j][s]
//   A_s = Kt + c_s (Mr - j Mi) + sum_t d_{s,t} T_t       (T_t on the boundary overlay rows only)
// x is system-interleaved (x[j*S + s]): one pass over the frequency-independent matrices serves all S
// systems and each gathered column is S contiguous values.  y is system-major (y[s*n + row]).
// Wide form (S > 8): one wavefront per row, lane l owns systems l + 64m (m < MS); a block covers 16
// consecutive rows (4 per wavefront) so each written cache line of y comes from one workgroup.
template <typename V, typename X, int MS, bool LOSSY>
__global__ __launch_bounds__(256) void k_spmv_sweep_w(SweepMatrix<V> M, int S, int R, const V *__restrict__ csys,
                                                      const V *__restrict__ dsys, const uint8_t *__restrict__ rowmask,
                                                      const X *__restrict__ x, X *__restrict__ y, int mode,
                                                      const int32_t *__restrict__ active) {
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int lane = threadIdx.x & 63;
  V cs[MS];
  bool on[MS];
#pragma unroll
  for (int m = 0; m < MS; m++) {
    const int s = lane + 64 * m;
    on[m] = s < S && (active == nullptr || active[s] != 0);
    cs[m] = s < S ? csys[s] : (V)0;
  }
  for (int rr = 0; rr < 4; rr++) {
    const int64_t row = (int64_t)blockIdx.x * 16 + wave * 4 + rr;
    if (row >= M.n) break;
    const int beg = M.rowptr[row], end = M.rowptr[row + 1];
    V ax[MS], ay[MS];
#pragma unroll
    for (int m = 0; m < MS; m++) ax[m] = ay[m] = (V)0;
    for (int k = beg; k < end; k++) {
      const int64_t j = M.colind[k];
      const V kv = M.Kt[k], mr = M.Mr[k];
      V mi = (V)0;
      if constexpr (LOSSY) mi = M.Mi[k];
#pragma unroll
      for (int m = 0; m < MS; m++) {
        if (on[m]) {
          const X xv = x[j * S + lane + 64 * m];
          const V ar = kv + cs[m] * mr;
          if constexpr (LOSSY) {
            const V ai = -cs[m] * mi;
            ax[m] += ar * xv.x - ai * xv.y;
            ay[m] += ar * xv.y + ai * xv.x;
          } else {
            ax[m] += ar * xv.x;
            ay[m] += ar * xv.y;
          }
        }
      }
    }
    if (M.nterm > 0) {
      for (int e = M.ov_ptr[row]; e < M.ov_ptr[row + 1]; e++) {
        const int64_t j = M.ov_col[e];
#pragma unroll
        for (int m = 0; m < MS; m++) {
          if (on[m]) {
            const int s = lane + 64 * m;
            V cr = (V)0, ci = (V)0;
            for (int t = 0; t < M.nterm; t++) {
              const V tv = M.ov_val[(int64_t)t * M.nov + e];
              cr += dsys[((int64_t)s * M.nterm + t) * 2] * tv;
              ci += dsys[((int64_t)s * M.nterm + t) * 2 + 1] * tv;
            }
            const X xv = x[j * S + s];
            ax[m] += cr * xv.x - ci * xv.y;
            ay[m] += cr * xv.y + ci * xv.x;
          }
        }
      }
    }
#pragma unroll
    for (int m = 0; m < MS; m++) {
      if (on[m]) {
        const int s = lane + 64 * m;
        const bool free_row = rowmask[(int64_t)(s % R) * M.n + row] != 0;
        X o;
        o.x = free_row ? (mode ? -ax[m] : ax[m]) : (V)0;
        o.y = free_row ? (mode ? -ay[m] : ay[m]) : (V)0;
        y[(int64_t)s * M.n + row] = o;
      }
    }
  }
}

// Tiled form (S > 8): a workgroup covers 16 rows x one tile of 16 systems, lane (row r, system s) walks row r
// for system s, so each gathered column is one 128-byte segment of x.  Workgroups are numbered tile-fastest:
// with 8 tiles (S = 128) workgroup b runs on XCD b % 8 = its tile, so each XCD streams one tile's rows in
// order and keeps that tile's window of x in its own L2 (the matrix is shared through the MALL).
template <typename V, typename X, bool LOSSY>
__global__ __launch_bounds__(256) void k_spmv_sweep_t(SweepMatrix<V> M, int S, int R, const V *__restrict__ csys,
                                                      const V *__restrict__ dsys, const uint8_t *__restrict__ rowmask,
                                                      const X *__restrict__ x, X *__restrict__ y, int mode,
                                                      const int32_t *__restrict__ active) {
  const int ntile = (S + 15) >> 4;
  const int64_t b = blockIdx.x;
  const int tile = (int)(b % ntile);
  const int64_t row = (b / ntile) * 16 + (threadIdx.x >> 4);
  const int s = tile * 16 + (threadIdx.x & 15);
  if (row >= M.n) return;
  const bool on = s < S && (active == nullptr || active[s] != 0);
  if (!on) return;
  const V cs = csys[s];
  V ax = (V)0, ay = (V)0;
  const int beg = M.rowptr[row], end = M.rowptr[row + 1];
#pragma unroll 4
  for (int k = beg; k < end; k++) {
    const int64_t j = M.colind[k];
    const V ar = M.Kt[k] + cs * M.Mr[k];
    const X xv = x[j * S + s];
    if constexpr (LOSSY) {
      const V ai = -cs * M.Mi[k];
      ax += ar * xv.x - ai * xv.y;
      ay += ar * xv.y + ai * xv.x;
    } else {
      ax += ar * xv.x;
      ay += ar * xv.y;
    }
  }
  if (M.nterm > 0) {
    for (int e = M.ov_ptr[row]; e < M.ov_ptr[row + 1]; e++) {
      const int64_t j = M.ov_col[e];
      V cr = (V)0, ci = (V)0;
      for (int t = 0; t < M.nterm; t++) {
        const V tv = M.ov_val[(int64_t)t * M.nov + e];
        cr += dsys[((int64_t)s * M.nterm + t) * 2] * tv;
        ci += dsys[((int64_t)s * M.nterm + t) * 2 + 1] * tv;
      }
      const X xv = x[j * S + s];
      ax += cr * xv.x - ci * xv.y;
      ay += cr * xv.y + ci * xv.x;
    }
  }
  const bool free_row = rowmask[(int64_t)(s % R) * M.n + row] != 0;
  X o;
  o.x = free_row ? (mode ? -ax : ax) : (V)0;
  o.y = free_row ? (mode ? -ay : ay) : (V)0;
  y[(int64_t)s * M.n + row] = o;
}

// Paired form (S > 8, S even): a workgroup covers 16 rows x one tile of 32 systems; lane (row r, l) owns
// systems s, s+1 (s = 32*tile + 2l), so a gathered column is one 16-byte load per lane (FP32) and one
// 256-byte segment per row.  The workgroup's matrix entries (column, Kt, Mr[, Mi]) are staged in LDS
// first, so the row loop issues one global load per two systems.  Workgroups are numbered tile-fastest.
template <typename X>
struct Pair;
template <>
struct Pair<float2> {
  using T = float4;
};
template <>
struct Pair<double2> {
  struct T {
    double x, y, z, w;
  };
};

template <typename V, typename X, bool LOSSY>
__global__ __launch_bounds__(256) void k_spmv_sweep_p(SweepMatrix<V> M, int S, int R, const V *__restrict__ csys,
                                                      const V *__restrict__ dsys, const uint8_t *__restrict__ rowmask,
                                                      const X *__restrict__ x, X *__restrict__ y, int mode,
                                                      const int32_t *__restrict__ active) {
  using P2 = typename Pair<X>::T;
  extern __shared__ __align__(16) unsigned char smem[];
  const int ntile = (S + 31) >> 5;
  const int64_t b = blockIdx.x;
  const int tile = (int)(b % ntile);
  const int64_t r0 = (b / ntile) * 16;
  const int64_t rend = r0 + 16 < M.n ? r0 + 16 : M.n;
  const int e0 = M.rowptr[r0], ne = M.rowptr[rend] - e0;
  V *lk = (V *)smem;
  V *lm = lk + M.blk16_max;
  V *lmi = lm + M.blk16_max;
  int32_t *lcol = (int32_t *)(lmi + (LOSSY ? M.blk16_max : 0));
  for (int e = threadIdx.x; e < ne; e += blockDim.x) {
    lcol[e] = M.colind[e0 + e];
    lk[e] = M.Kt[e0 + e];
    lm[e] = M.Mr[e0 + e];
    if constexpr (LOSSY) lmi[e] = M.Mi[e0 + e];
  }
  __syncthreads();
  const int64_t row = r0 + (threadIdx.x >> 4);
  const int s = tile * 32 + 2 * (threadIdx.x & 15);
  if (row >= M.n || s >= S) return;
  const bool on0 = active == nullptr || active[s] != 0, on1 = active == nullptr || active[s + 1] != 0;
  if (!on0 && !on1) return;
  const V c0 = csys[s], c1 = csys[s + 1];
  V a0x = (V)0, a0y = (V)0, a1x = (V)0, a1y = (V)0;
  const int kb = M.rowptr[row] - e0, ke = M.rowptr[row + 1] - e0;
#pragma unroll 4
  for (int k = kb; k < ke; k++) {
    const int64_t j = lcol[k];
    const V kv = lk[k], mr = lm[k];
    const P2 xv = *(const P2 *)(x + j * S + s);
    const V r0v = kv + c0 * mr, r1v = kv + c1 * mr;
    if constexpr (LOSSY) {
      const V mi = lmi[k];
      const V i0 = -c0 * mi, i1 = -c1 * mi;
      a0x += r0v * xv.x - i0 * xv.y;
      a0y += r0v * xv.y + i0 * xv.x;
      a1x += r1v * xv.z - i1 * xv.w;
      a1y += r1v * xv.w + i1 * xv.z;
    } else {
      a0x += r0v * xv.x;
      a0y += r0v * xv.y;
      a1x += r1v * xv.z;
      a1y += r1v * xv.w;
    }
  }
  if (M.nterm > 0) {
    for (int e = M.ov_ptr[row]; e < M.ov_ptr[row + 1]; e++) {
      const int64_t j = M.ov_col[e];
      V d0r = (V)0, d0i = (V)0, d1r = (V)0, d1i = (V)0;
      for (int t = 0; t < M.nterm; t++) {
        const V tv = M.ov_val[(int64_t)t * M.nov + e];
        d0r += dsys[((int64_t)s * M.nterm + t) * 2] * tv;
        d0i += dsys[((int64_t)s * M.nterm + t) * 2 + 1] * tv;
        d1r += dsys[((int64_t)(s + 1) * M.nterm + t) * 2] * tv;
        d1i += dsys[((int64_t)(s + 1) * M.nterm + t) * 2 + 1] * tv;
      }
      const P2 xv = *(const P2 *)(x + j * S + s);
      a0x += d0r * xv.x - d0i * xv.y;
      a0y += d0r * xv.y + d0i * xv.x;
      a1x += d1r * xv.z - d1i * xv.w;
      a1y += d1r * xv.w + d1i * xv.z;
    }
  }
  if (on0) {
    const bool fr = rowmask[(int64_t)(s % R) * M.n + row] != 0;
    X o;
    o.x = fr ? (mode ? -a0x : a0x) : (V)0;
    o.y = fr ? (mode ? -a0y : a0y) : (V)0;
    y[(int64_t)s * M.n + row] = o;
  }
  if (on1) {
    const bool fr = rowmask[(int64_t)((s + 1) % R) * M.n + row] != 0;
    X o;
    o.x = fr ? (mode ? -a1x : a1x) : (V)0;
    o.y = fr ? (mode ? -a1y : a1y) : (V)0;
    y[(int64_t)(s + 1) * M.n + row] = o;
  }
}

// Narrow form (S <= 8): LPR lanes split a row's entries, each lane carries all S systems.
template <typename V, typename X, int LPR, int SP, bool LOSSY>
__global__ __launch_bounds__(256) void k_spmv_sweep_n(SweepMatrix<V> M, int S, int R, const V *__restrict__ csys,
                                                      const V *__restrict__ dsys, const uint8_t *__restrict__ rowmask,
                                                      const X *__restrict__ x, X *__restrict__ y, int mode,
                                                      const int32_t *__restrict__ active) {
  const int sub = threadIdx.x % LPR;
  const int64_t row = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / LPR;
  if (row >= M.n) return;
  V cs[SP], ax[SP], ay[SP];
#pragma unroll
  for (int s = 0; s < SP; s++) {
    cs[s] = s < S ? csys[s] : (V)0;
    ax[s] = ay[s] = (V)0;
  }
  for (int k = M.rowptr[row] + sub; k < M.rowptr[row + 1]; k += LPR) {
    const int64_t j = M.colind[k];
    const V kv = M.Kt[k], mr = M.Mr[k];
    V mi = (V)0;
    if constexpr (LOSSY) mi = M.Mi[k];
#pragma unroll
    for (int s = 0; s < SP; s++) {
      if (s < S) {
        const X xv = x[j * S + s];
        const V ar = kv + cs[s] * mr;
        if constexpr (LOSSY) {
          const V ai = -cs[s] * mi;
          ax[s] += ar * xv.x - ai * xv.y;
          ay[s] += ar * xv.y + ai * xv.x;
        } else {
          ax[s] += ar * xv.x;
          ay[s] += ar * xv.y;
        }
      }
    }
  }
  if (M.nterm > 0) {
    for (int e = M.ov_ptr[row] + sub; e < M.ov_ptr[row + 1]; e += LPR) {
      const int64_t j = M.ov_col[e];
#pragma unroll
      for (int s = 0; s < SP; s++) {
        if (s < S) {
          V cr = (V)0, ci = (V)0;
          for (int t = 0; t < M.nterm; t++) {
            const V tv = M.ov_val[(int64_t)t * M.nov + e];
            cr += dsys[((int64_t)s * M.nterm + t) * 2] * tv;
            ci += dsys[((int64_t)s * M.nterm + t) * 2 + 1] * tv;
          }
          const X xv = x[j * S + s];
          ax[s] += cr * xv.x - ci * xv.y;
          ay[s] += cr * xv.y + ci * xv.x;
        }
      }
    }
  }
#pragma unroll
  for (int s = 0; s < SP; s++) {
#pragma unroll
    for (int o = LPR / 2; o > 0; o >>= 1) {
      ax[s] += __shfl_xor(ax[s], o, LPR);
      ay[s] += __shfl_xor(ay[s], o, LPR);
    }
  }
  if (sub == 0) {
#pragma unroll
    for (int s = 0; s < SP; s++) {
      if (s < S && (active == nullptr || active[s] != 0)) {
        const bool free_row = rowmask[(int64_t)(s % R) * M.n + row] != 0;
        X o;
        o.x = free_row ? (mode ? -ax[s] : ax[s]) : (V)0;
        o.y = free_row ? (mode ? -ay[s] : ay[s]) : (V)0;
        y[(int64_t)s * M.n + row] = o;
      }
    }
  }
}

template <typename V, typename X, bool LOSSY>
static void spmv_sweep_l(hipStream_t st, const SweepMatrix<V> &M, int S, int R, const V *csys, const V *dsys,
                         const uint8_t *rowmask, const X *x, X *y, int mode, const int32_t *active, int lpr, int form) {
  if (S > 8 && form == 0) form = S % 2 == 0 && spmv_pair_enabled() ? 1 : spmv_wide_enabled() ? 3 : 2;
  if (S > 8 && form == 1) {
    if (S > 256) throw std::runtime_error("sweep spmv: at most 256 systems per batch");
    if (M.blk16_max <= 0) throw std::runtime_error("sweep spmv: blk16_max not set");
    const int64_t nb = ((M.n + 15) / 16) * ((S + 31) / 32);
    const size_t lds = (size_t)M.blk16_max * ((LOSSY ? 3 : 2) * sizeof(V) + 4);
    if (lds > 64 * 1024) throw std::runtime_error("sweep spmv: rows too long for the staged form");
    hipLaunchKernelGGL((k_spmv_sweep_p<V, X, LOSSY>), dim3((unsigned)nb), dim3(256), lds, st, M, S, R, csys, dsys,
                       rowmask, x, y, mode, active);
  } else if (S > 8 && form == 2) {
    const int64_t nb = ((M.n + 15) / 16) * ((S + 15) / 16);
    if (S > 256) throw std::runtime_error("sweep spmv: at most 256 systems per batch");
    hipLaunchKernelGGL((k_spmv_sweep_t<V, X, LOSSY>), dim3((unsigned)nb), dim3(256), 0, st, M, S, R, csys, dsys,
                       rowmask, x, y, mode, active);
  } else if (S > 8) {
    dim3 grid((unsigned)((M.n + 15) / 16));
    if (S <= 64)
      hipLaunchKernelGGL((k_spmv_sweep_w<V, X, 1, LOSSY>), grid, dim3(256), 0, st, M, S, R, csys, dsys, rowmask, x,
                         y, mode, active);
    else if (S <= 128)
      hipLaunchKernelGGL((k_spmv_sweep_w<V, X, 2, LOSSY>), grid, dim3(256), 0, st, M, S, R, csys, dsys, rowmask, x,
                         y, mode, active);
    else if (S <= 256)
      hipLaunchKernelGGL((k_spmv_sweep_w<V, X, 4, LOSSY>), grid, dim3(256), 0, st, M, S, R, csys, dsys, rowmask, x,
                         y, mode, active);
    else
      throw std::runtime_error("sweep spmv: at most 256 systems per batch");
  } else {
    const int L = lpr <= 8 ? 8 : lpr <= 16 ? 16 : 32;
    dim3 grid((unsigned)((M.n * L + 255) / 256));
#define OPEM_SPMV_N(LL, SS)                                                                                   \
  hipLaunchKernelGGL((k_spmv_sweep_n<V, X, LL, SS, LOSSY>), grid, dim3(256), 0, st, M, S, R, csys, dsys, rowmask, \
                     x, y, mode, active)
#define OPEM_SPMV_NS(LL)        \
  if (S <= 1)                   \
    OPEM_SPMV_N(LL, 1);         \
  else if (S <= 2)              \
    OPEM_SPMV_N(LL, 2);         \
  else if (S <= 4)              \
    OPEM_SPMV_N(LL, 4);         \
  else                          \
    OPEM_SPMV_N(LL, 8);
    if (L == 8) {
      OPEM_SPMV_NS(8)
    } else if (L == 16) {
      OPEM_SPMV_NS(16)
    } else {
      OPEM_SPMV_NS(32)
    }
#undef OPEM_SPMV_NS
#undef OPEM_SPMV_N
  }
  HIP_LAUNCH_CHECK();
}

template <typename V, typename X>
void launch_spmv_sweep(hipStream_t st, const SweepMatrix<V> &M, int S, int R, const V *csys, const V *dsys,
                       const uint8_t *rowmask, const X *x, X *y, int mode, const int32_t *active, int lpr, int form) {
  if (form < 0 || form > 3) throw std::runtime_error("sweep spmv: form must be 0..3");
  if (M.Mi)
    spmv_sweep_l<V, X, true>(st, M, S, R, csys, dsys, rowmask, x, y, mode, active, lpr, form);
  else
    spmv_sweep_l<V, X, false>(st, M, S, R, csys, dsys, rowmask, x, y, mode, active, lpr, form);
}
template void launch_spmv_sweep<double, double2>(hipStream_t, const SweepMatrix<double> &, int, int, const double *,
                                                 const double *, const uint8_t *, const double2 *, double2 *, int,
                                                 const int32_t *, int, int);
template void launch_spmv_sweep<float, float2>(hipStream_t, const SweepMatrix<float> &, int, int, const float *,
                                               const float *, const uint8_t *, const float2 *, float2 *, int,
                                               const int32_t *, int, int);

int spmv_tune_level() {
  static int v = -1;
  if (v < 0) {
    const char *e = getenv("OPEM_SPMV_TUNE");
    // a form forced by OPEM_SPMV_PAIR / OPEM_SPMV_WIDE stays forced unless tuning is asked for explicitly
    const bool forced = getenv("OPEM_SPMV_PAIR") != nullptr || getenv("OPEM_SPMV_WIDE") != nullptr;
    v = e ? atoi(e) : forced ? 0 : 1;
    if (v < 0 || v > 2) v = 1;
  }
  return v;
}

__global__ __launch_bounds__(256) void k_hash64(int64_t count, const uint64_t *__restrict__ a,
                                                unsigned long long *__restrict__ out) {
  unsigned long long acc = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
    acc += (unsigned long long)a[i] * (2ull * (unsigned long long)i + 1ull);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if ((threadIdx.x & 63) == 0) atomicAdd(out, acc);
}

void launch_hash64(hipStream_t st, int64_t count, const uint64_t *a, unsigned long long *out) {
  if (count <= 0) return;
  int64_t blocks = (count + 255) / 256;
  if (blocks > 2048) blocks = 2048;
  hipLaunchKernelGGL(k_hash64, dim3((unsigned)blocks), dim3(256), 0, st, count, a, out);
  HIP_LAUNCH_CHECK();
}

// Kt = sum_a wk[a] K[a], Mr = sum_a wr[a] M[a], Mi = sum_a wi[a] M[a] (FP64, plus FP32 copies if f32 != null)
__global__ void k_base_matrices(int64_t nnz, int nslot, const double *__restrict__ K, const double *__restrict__ M,
                                const double *__restrict__ w, double *__restrict__ out64, float *__restrict__ out32,
                                int lossy) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nnz; i += (int64_t)gridDim.x * blockDim.x) {
    double kt = 0.0, mr = 0.0, mi = 0.0;
    for (int a = 0; a < nslot; a++) {
      const double kv = K[(int64_t)a * nnz + i], mv = M[(int64_t)a * nnz + i];
      kt += w[a * 3] * kv;
      mr += w[a * 3 + 1] * mv;
      mi += w[a * 3 + 2] * mv;
    }
    out64[i] = kt;
    out64[nnz + i] = mr;
    if (lossy) out64[2 * nnz + i] = mi;
    if (out32) {
      out32[i] = (float)kt;
      out32[nnz + i] = (float)mr;
      if (lossy) out32[2 * nnz + i] = (float)mi;
    }
  }
}

void launch_base_matrices(hipStream_t st, int64_t nnz, int nslot, const double *K, const double *M, const double *w,
                          double *out64, float *out32, int lossy) {
  int64_t bl = (nnz + 255) / 256;
  if (bl > 65536) bl = 65536;
  hipLaunchKernelGGL(k_base_matrices, dim3((unsigned)bl), dim3(256), 0, st, nnz, nslot, K, M, w, out64, out32, lossy);
  HIP_LAUNCH_CHECK();
}

// out[i*S + s] = in[s*n + i] (to_il) or out[s*n + i] = in[i*S + s], through 32 x 32 LDS tiles
template <typename X, bool TO_IL>
__global__ __launch_bounds__(256) void k_transpose(int64_t n, int S, const X *__restrict__ in, X *__restrict__ out) {
  __shared__ X tile[32][33];
  const int64_t i0 = (int64_t)blockIdx.x * 32;
  const int s0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  if (TO_IL) {
    for (int k = ty; k < 32; k += 8) {  // read rows s, columns i (coalesced over i)
      const int s = s0 + k;
      const int64_t i = i0 + tx;
      if (s < S && i < n) tile[k][tx] = in[(int64_t)s * n + i];
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {  // write rows i, columns s (coalesced over s)
      const int64_t i = i0 + k;
      const int s = s0 + tx;
      if (s < S && i < n) out[i * S + s] = tile[tx][k];
    }
  } else {
    for (int k = ty; k < 32; k += 8) {
      const int64_t i = i0 + k;
      const int s = s0 + tx;
      if (s < S && i < n) tile[tx][k] = in[i * S + s];
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
      const int s = s0 + k;
      const int64_t i = i0 + tx;
      if (s < S && i < n) out[(int64_t)s * n + i] = tile[k][tx];
    }
  }
}

template <typename X>
void launch_transpose(hipStream_t st, int64_t n, int S, const X *in, X *out, bool to_il) {
  dim3 grid((unsigned)((n + 31) / 32), (unsigned)((S + 31) / 32));
  if (to_il)
    hipLaunchKernelGGL((k_transpose<X, true>), grid, dim3(256), 0, st, n, S, in, out);
  else
    hipLaunchKernelGGL((k_transpose<X, false>), grid, dim3(256), 0, st, n, S, in, out);
  HIP_LAUNCH_CHECK();
}
template void launch_transpose<double2>(hipStream_t, int64_t, int, const double2 *, double2 *, bool);
template void launch_transpose<float2>(hipStream_t, int64_t, int, const float2 *, float2 *, bool);

// x[i*S + s] += z[i*S + s] for active s (both system-interleaved)
template <typename X, typename Z>
__global__ void k_axpy_il(int64_t n, int S, const Z *__restrict__ z, X *__restrict__ x,
                          const int32_t *__restrict__ active) {
  const int64_t tot = n * S;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (int64_t)gridDim.x * blockDim.x) {
    const int s = (int)(e % S);
    if (active && !active[s]) continue;
    const double2 a = to_d2(x[e]), b = to_d2(z[e]);
    x[e] = from_d2<X>(make_double2(a.x + b.x, a.y + b.y));
  }
}
template <typename X, typename Z>
void launch_axpy_il(hipStream_t st, int S, int64_t n, const Z *z, X *x, const int32_t *active) {
  int64_t bl = (n * S + 255) / 256;
  if (bl > 65536) bl = 65536;
  hipLaunchKernelGGL((k_axpy_il<X, Z>), dim3((unsigned)bl), dim3(256), 0, st, n, S, z, x, active);
  HIP_LAUNCH_CHECK();
}
template void launch_axpy_il<double2, double2>(hipStream_t, int, int64_t, const double2 *, double2 *,
                                               const int32_t *);
template void launch_axpy_il<double2, float2>(hipStream_t, int, int64_t, const float2 *, double2 *, const int32_t *);

// ------------------------------------------------------------------ polynomial preconditioner steps
// Step i of p(B) v (Loe-Morgan): u += c prod; prod -= c t, with c = coef[s*ldc + 2i..] (complex, per system);
// t == null: last step (u += c prod only); first: u = c prod.
template <typename X>
__global__ void k_poly_step(int64_t n, const double *__restrict__ coef, int ldc, X *__restrict__ prod,
                            const X *__restrict__ t, X *__restrict__ u, int first, const int32_t *__restrict__ active) {
  const int s = blockIdx.y;
  if (active && !active[s]) return;
  const double cr = coef[(int64_t)s * ldc], ci = coef[(int64_t)s * ldc + 1];
  X *ps = prod + (int64_t)s * n;
  X *us = u + (int64_t)s * n;
  const X *ts = t ? t + (int64_t)s * n : nullptr;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double2 p = to_d2(ps[i]);
    double2 o = first ? make_double2(0.0, 0.0) : to_d2(us[i]);
    o.x += cr * p.x - ci * p.y;
    o.y += cr * p.y + ci * p.x;
    us[i] = from_d2<X>(o);
    if (ts) {
      const double2 q = to_d2(ts[i]);
      ps[i] = from_d2<X>(make_double2(p.x - (cr * q.x - ci * q.y), p.y - (cr * q.y + ci * q.x)));
    }
  }
}

template <typename X>
void launch_poly_step(hipStream_t st, int S, int64_t n, const double *coef, int ldc, X *prod, const X *t, X *u,
                      bool first, const int32_t *active) {
  int64_t bl = (n + 255) / 256;
  if (bl > 4096) bl = 4096;
  hipLaunchKernelGGL((k_poly_step<X>), dim3((unsigned)bl, S), dim3(256), 0, st, n, coef, ldc, prod, t, u,
                     first ? 1 : 0, active);
  HIP_LAUNCH_CHECK();
}
template void launch_poly_step<double2>(hipStream_t, int, int64_t, const double *, int, double2 *, const double2 *,
                                        double2 *, bool, const int32_t *);
template void launch_poly_step<float2>(hipStream_t, int, int64_t, const double *, int, float2 *, const float2 *,
                                       float2 *, bool, const int32_t *);

// out[s][i] = in[s*ldin + i] for active s
template <typename X>
__global__ void k_copy_sys(int64_t n, const X *__restrict__ in, int64_t ldin, X *__restrict__ out, int64_t ldout,
                           const int32_t *__restrict__ active) {
  const int s = blockIdx.y;
  if (active && !active[s]) return;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[(int64_t)s * ldout + i] = in[(int64_t)s * ldin + i];
}

template <typename X>
void launch_copy_sys(hipStream_t st, int S, int64_t n, const X *in, int64_t ldin, X *out, int64_t ldout,
                     const int32_t *active) {
  int64_t bl = (n + 255) / 256;
  if (bl > 4096) bl = 4096;
  hipLaunchKernelGGL((k_copy_sys<X>), dim3((unsigned)bl, S), dim3(256), 0, st, n, in, ldin, out, ldout, active);
  HIP_LAUNCH_CHECK();
}
template void launch_copy_sys<double2>(hipStream_t, int, int64_t, const double2 *, int64_t, double2 *, int64_t,
                                       const int32_t *);
template void launch_copy_sys<float2>(hipStream_t, int, int64_t, const float2 *, int64_t, float2 *, int64_t,
                                      const int32_t *);

// ------------------------------------------------------------------ Krylov helpers
constexpr int kMaxK = kMaxBasis;  // basis vectors per reduction (restart + 1)

template <typename X, int E>
__global__ __launch_bounds__(kRedThreads) void k_mdot(int64_t n, int k, const X *__restrict__ Vb, int64_t ldv_sys,
                                                      const X *__restrict__ w, int64_t ldw_sys,
                                                      double *__restrict__ partial, int nblk,
                                                      const int32_t *__restrict__ active) {
  const int s = blockIdx.y;
  if (active && !active[s]) return;
  const int64_t base = (int64_t)blockIdx.x * (kRedThreads * E) + threadIdx.x;
  const X *ws = w + s * ldw_sys;
  const X *Vs = Vb + s * ldv_sys;
  double2 wl[E];
#pragma unroll
  for (int e = 0; e < E; e++) {
    const int64_t i = base + e * kRedThreads;
    wl[e] = i < n ? to_d2(ws[i]) : make_double2(0.0, 0.0);
  }
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  // four basis vectors per pass; their wave sums by a reduce-scatter butterfly: after the xor-32 and
  // xor-16 steps each 16-lane quarter holds one vector's partial sums, finished by xor 8, 4, 2, 1
  const bool h32 = (lane & 32) != 0, h16 = (lane & 16) != 0;
  for (int v0 = 0; v0 < k; v0 += 4) {
    double ax[4], ay[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      ax[q] = ay[q] = 0.0;
      if (v0 + q < k) {
#pragma unroll
        for (int e = 0; e < E; e++) {
          const int64_t i = base + e * kRedThreads;
          if (i < n) {
            const double2 a = to_d2(Vs[(int64_t)(v0 + q) * n + i]);
            ax[q] += a.x * wl[e].x + a.y * wl[e].y;
            ay[q] += a.x * wl[e].y - a.y * wl[e].x;
          }
        }
      }
    }
    double bx[2], by[2];
#pragma unroll
    for (int t = 0; t < 2; t++) {
      const double kx = h32 ? ax[t + 2] : ax[t], ky = h32 ? ay[t + 2] : ay[t];
      const double sx = h32 ? ax[t] : ax[t + 2], sy = h32 ? ay[t] : ay[t + 2];
      bx[t] = kx + __shfl_xor(sx, 32, 64);
      by[t] = ky + __shfl_xor(sy, 32, 64);
    }
    double cx = (h16 ? bx[1] : bx[0]) + __shfl_xor(h16 ? bx[0] : bx[1], 16, 64);
    double cy = (h16 ? by[1] : by[0]) + __shfl_xor(h16 ? by[0] : by[1], 16, 64);
#pragma unroll
    for (int o = 8; o > 0; o >>= 1) {
      cx += __shfl_xor(cx, o, 64);
      cy += __shfl_xor(cy, o, 64);
    }
    const int q = (h32 ? 2 : 0) + (h16 ? 1 : 0);
    if ((lane & 15) == 0 && v0 + q < k) {
      double *p = partial + (((int64_t)s * nblk + blockIdx.x * (kRedThreads / 64) + wave) * kMaxK + v0 + q) * 2;
      p[0] = cx;
      p[1] = cy;
    }
  }
}

constexpr int kMdotElems = 8;  // elements per thread (4 and 16 measured slower at 128 systems x 76 vectors)

static int mdot_grid(int64_t n) {
  const int64_t c = (int64_t)kRedThreads * kMdotElems;
  const int64_t b = (n + c - 1) / c;
  return (int)(b < 1 ? 1 : b);
}

int mdot_blocks(int64_t n) { return mdot_grid(n) * (kRedThreads / 64); }

template <typename X>
void launch_mdot(hipStream_t st, int S, int64_t n, int k, const X *Vb, int64_t ldv_sys, const X *w,
                 int64_t ldw_sys, double *partial, int nblk, const int32_t *active) {
  if (k > kMaxK) throw std::runtime_error("restart too large");
  if (nblk != mdot_blocks(n)) throw std::runtime_error("mdot: partial block count mismatch");
  hipLaunchKernelGGL((k_mdot<X, kMdotElems>), dim3(mdot_grid(n), S), dim3(kRedThreads), 0, st, n, k, Vb, ldv_sys, w,
                     ldw_sys, partial, nblk, active);
  HIP_LAUNCH_CHECK();
}
template void launch_mdot<double2>(hipStream_t, int, int64_t, int, const double2 *, int64_t, const double2 *,
                                   int64_t, double *, int, const int32_t *);
template void launch_mdot<float2>(hipStream_t, int, int64_t, int, const float2 *, int64_t, const float2 *, int64_t,
                                  double *, int, const int32_t *);

template <typename X>
__global__ __launch_bounds__(kRedThreads) void k_mupdate(int64_t n, int k, const X *__restrict__ Vb, int64_t ldv_sys,
                                                         X *__restrict__ w, int64_t ldw_sys,
                                                         const double *__restrict__ h, int ldh,
                                                         double *__restrict__ part2, int nblk,
                                                         const int32_t *__restrict__ active) {
  const int s = blockIdx.y;
  if (active && !active[s]) return;
  __shared__ double2 hs[kMaxK];
  __shared__ double red[kRedThreads / 64];
  for (int v = threadIdx.x; v < k; v += kRedThreads)
    hs[v] = make_double2(h[((int64_t)s * ldh + v) * 2], h[((int64_t)s * ldh + v) * 2 + 1]);
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kChunk + threadIdx.x;
  X *ws = w + s * ldw_sys;
  const X *Vs = Vb + s * ldv_sys;
  double nrm = 0.0;
#pragma unroll
  for (int e = 0; e < kRedElems; e++) {
    const int64_t i = base + e * kRedThreads;
    if (i < n) {
      double2 acc = to_d2(ws[i]);
      for (int v = 0; v < k; v++) {
        const double2 a = to_d2(Vs[(int64_t)v * n + i]);
        const double2 c = cmul(a, hs[v]);
        acc.x -= c.x;
        acc.y -= c.y;
      }
      const X out = from_d2<X>(acc);
      ws[i] = out;
      const double2 o = to_d2(out);
      nrm += o.x * o.x + o.y * o.y;
    }
  }
  if (part2) {
    nrm = wave_sum(nrm);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = nrm;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0.0;
      for (int q = 0; q < kRedThreads / 64; q++) t += red[q];
      part2[(int64_t)s * nblk + blockIdx.x] = t;
    }
  }
}

template <typename X>
void launch_mupdate(hipStream_t st, int S, int64_t n, int k, const X *Vb, int64_t ldv_sys, X *w, int64_t ldw_sys,
                    const double *h, int ldh, double *part2, int nblk, const int32_t *active) {
  hipLaunchKernelGGL((k_mupdate<X>), dim3(nblk, S), dim3(kRedThreads), 0, st, n, k, Vb, ldv_sys, w, ldw_sys, h, ldh,
                     part2, nblk, active);
  HIP_LAUNCH_CHECK();
}
template void launch_mupdate<double2>(hipStream_t, int, int64_t, int, const double2 *, int64_t, double2 *, int64_t,
                                      const double *, int, double *, int, const int32_t *);
template void launch_mupdate<float2>(hipStream_t, int, int64_t, int, const float2 *, int64_t, float2 *, int64_t,
                                     const double *, int, double *, int, const int32_t *);

// ------------------------------------------------------------------ s-step (block) Arnoldi, s = 2..4
// The s new vectors z_c (c < NS) live in one buffer, z_c of system s at Z + (c*S + s)*n.
// Projection in one pass over V: partial[s][blk][v][c] = sum_{chunk} conj(V[s][v][i]) z_c[s][i] (complex), one
// partial per wavefront; each pass takes VB = 4/NS basis vectors (4 complex sums) and finishes them with the
// reduce-scatter butterfly of k_mdot (NS = 3 leaves one of the four slots empty).
template <typename X, int E, int NS>
__global__ __launch_bounds__(kRedThreads) void k_mdot_s(int64_t n, int k, const X *__restrict__ Vb, int64_t ldv_sys,
                                                        const X *__restrict__ Z, double *__restrict__ partial,
                                                        int nblk, const int32_t *__restrict__ active) {
  constexpr int VB = NS >= 3 ? 1 : 4 / NS;
  // basis groups whose loads are issued together before any of them is used (FP32 basis: two groups, so a
  // lane has 2*VB*E loads in flight; FP64 keeps one group for registers)
  constexpr int U = std::is_same<X, float2>::value ? 2 : 1;
  const int s = blockIdx.y, S = gridDim.y;
  if (active && !active[s]) return;
  const int64_t base = (int64_t)blockIdx.x * (kRedThreads * E) + threadIdx.x;
  const X *Vs = Vb + s * ldv_sys;
  // Element offsets clamped into [0, n): lanes past the end read a valid entry and multiply it by z = 0, so
  // the loads carry no branch and are issued back to back.
  uint32_t off[E];
  X w[NS][E];
#pragma unroll
  for (int e = 0; e < E; e++) {
    const int64_t i = base + e * kRedThreads;
    off[e] = (uint32_t)(i < n ? i : n - 1);
  }
#pragma unroll
  for (int c = 0; c < NS; c++) {
    const X *zc = Z + ((int64_t)c * S + s) * n;
#pragma unroll
    for (int e = 0; e < E; e++) {
      X t = zc[off[e]];
      if (base + e * kRedThreads >= n) t.x = t.y = 0;
      w[c][e] = t;
    }
  }
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const bool h32 = (lane & 32) != 0, h16 = (lane & 16) != 0;
  for (int v0 = 0; v0 < k; v0 += U * VB) {
    X a[U][VB][E];
#pragma unroll
    for (int u = 0; u < U; u++)
#pragma unroll
      for (int dv = 0; dv < VB; dv++) {
        const int v = min(v0 + u * VB + dv, k - 1);  // the tail re-reads the last vector; its sums are dropped
        const X *vp = Vs + (int64_t)v * n;
#pragma unroll
        for (int e = 0; e < E; e++) a[u][dv][e] = vp[off[e]];
      }
    if constexpr (U > 1) __builtin_amdgcn_sched_barrier(0);  // keep every group's loads ahead of the sums
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int vu = v0 + u * VB;
      double ax[4], ay[4];  // slot q = (v - vu) * NS + c
#pragma unroll
      for (int q = 0; q < 4; q++) ax[q] = ay[q] = 0.0;
#pragma unroll
      for (int dv = 0; dv < VB; dv++) {
        if constexpr (std::is_same<X, float2>::value) {
          // FP32 basis: the E products of a lane are summed in FP32, the lane sums in FP64
          float fx[NS], fy[NS];
#pragma unroll
          for (int c = 0; c < NS; c++) fx[c] = fy[c] = 0.f;
#pragma unroll
          for (int e = 0; e < E; e++) {
            const float2 t = a[u][dv][e];
#pragma unroll
            for (int c = 0; c < NS; c++) {
              fx[c] += t.x * w[c][e].x + t.y * w[c][e].y;
              fy[c] += t.x * w[c][e].y - t.y * w[c][e].x;
            }
          }
#pragma unroll
          for (int c = 0; c < NS; c++) {
            ax[dv * NS + c] = (double)fx[c];
            ay[dv * NS + c] = (double)fy[c];
          }
        } else {
#pragma unroll
          for (int e = 0; e < E; e++) {
            const double2 t = to_d2(a[u][dv][e]);
#pragma unroll
            for (int c = 0; c < NS; c++) {
              const double2 b = to_d2(w[c][e]);
              ax[dv * NS + c] += t.x * b.x + t.y * b.y;
              ay[dv * NS + c] += t.x * b.y - t.y * b.x;
            }
          }
        }
      }
      double bx[2], by[2];
#pragma unroll
      for (int t = 0; t < 2; t++) {
        const double kx = h32 ? ax[t + 2] : ax[t], ky = h32 ? ay[t + 2] : ay[t];
        const double sx = h32 ? ax[t] : ax[t + 2], sy = h32 ? ay[t] : ay[t + 2];
        bx[t] = kx + __shfl_xor(sx, 32, 64);
        by[t] = ky + __shfl_xor(sy, 32, 64);
      }
      double cx = (h16 ? bx[1] : bx[0]) + __shfl_xor(h16 ? bx[0] : bx[1], 16, 64);
      double cy = (h16 ? by[1] : by[0]) + __shfl_xor(h16 ? by[0] : by[1], 16, 64);
#pragma unroll
      for (int o = 8; o > 0; o >>= 1) {
        cx += __shfl_xor(cx, o, 64);
        cy += __shfl_xor(cy, o, 64);
      }
      const int q = (h32 ? 2 : 0) + (h16 ? 1 : 0);
      const int v = vu + q / NS, c = q % NS;
      if ((lane & 15) == 0 && q < VB * NS && v < k) {
        double *p =
            partial + ((((int64_t)s * nblk + blockIdx.x * (kRedThreads / 64) + wave) * kMaxK + v) * NS + c) * 2;
        p[0] = cx;
        p[1] = cy;
      }
    }
  }
}

template <typename X>
void launch_mdot_s(hipStream_t st, int ns, int S, int64_t n, int k, const X *Vb, int64_t ldv_sys, const X *Z,
                   double *partial, int nblk, const int32_t *active) {
  if (k > kMaxK) throw std::runtime_error("restart too large");
  if (nblk != mdot_blocks(n)) throw std::runtime_error("mdot_s: partial block count mismatch");
  if (n < 1 || n > INT32_MAX) throw std::runtime_error("mdot_s: vector length out of range");
  const dim3 grid(mdot_grid(n), S);
  if (ns == 2)
    hipLaunchKernelGGL((k_mdot_s<X, kMdotElems, 2>), grid, dim3(kRedThreads), 0, st, n, k, Vb, ldv_sys, Z, partial,
                       nblk, active);
  else if (ns == 3)
    hipLaunchKernelGGL((k_mdot_s<X, kMdotElems, 3>), grid, dim3(kRedThreads), 0, st, n, k, Vb, ldv_sys, Z, partial,
                       nblk, active);
  else if (ns == 4)
    hipLaunchKernelGGL((k_mdot_s<X, kMdotElems, 4>), grid, dim3(kRedThreads), 0, st, n, k, Vb, ldv_sys, Z, partial,
                       nblk, active);
  else
    throw std::runtime_error("mdot_s: block size must be 2..4");
  HIP_LAUNCH_CHECK();
}
template void launch_mdot_s<double2>(hipStream_t, int, int, int64_t, int, const double2 *, int64_t, const double2 *,
                                     double *, int, const int32_t *);
template void launch_mdot_s<float2>(hipStream_t, int, int, int64_t, int, const float2 *, int64_t, const float2 *,
                                    double *, int, const int32_t *);

// z_c -= sum_v V[v] C[v][c] (C[s][v][c] complex, row stride ldc blocks of NS), then the Gram matrix of the result,
// gram[s][blk][a][b] = z_a^H z_b (complex, NS x NS).
template <typename X, int NS>
__global__ __launch_bounds__(kRedThreads) void k_mupdate_s(int64_t n, int k, const X *__restrict__ Vb, int64_t ldv_sys,
                                                           X *__restrict__ Z, const double *__restrict__ C, int ldc,
                                                           double *__restrict__ gram, int nblk,
                                                           const int32_t *__restrict__ active) {
  const int s = blockIdx.y, S = gridDim.y;
  if (active && !active[s]) return;
  __shared__ double2 hs[NS][kMaxK];
  __shared__ double red[NS * NS * 2][kRedThreads / 64];
  for (int v = threadIdx.x; v < k; v += kRedThreads)
    for (int c = 0; c < NS; c++) {
      const double *cv = C + (((int64_t)s * ldc + v) * NS + c) * 2;
      hs[c][v] = make_double2(cv[0], cv[1]);
    }
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kChunk + threadIdx.x;
  const X *Vs = Vb + s * ldv_sys;
  double g[NS][NS][2];
#pragma unroll
  for (int a = 0; a < NS; a++)
#pragma unroll
    for (int b = 0; b < NS; b++) g[a][b][0] = g[a][b][1] = 0.0;
#pragma unroll
  for (int e = 0; e < kRedElems; e++) {
    const int64_t i = base + e * kRedThreads;
    if (i < n) {
      double2 p[NS];
#pragma unroll
      for (int c = 0; c < NS; c++) p[c] = to_d2(Z[((int64_t)c * S + s) * n + i]);
      // UV basis entries are loaded before any is used, so a lane keeps UV loads in flight
      constexpr int UV = 8;
      int v = 0;
      for (; v + UV <= k; v += UV) {
        X a[UV];
#pragma unroll
        for (int u = 0; u < UV; u++) a[u] = Vs[(int64_t)(v + u) * n + i];
#pragma unroll
        for (int u = 0; u < UV; u++)
#pragma unroll
          for (int c = 0; c < NS; c++) {
            const double2 t = cmul(to_d2(a[u]), hs[c][v + u]);
            p[c].x -= t.x;
            p[c].y -= t.y;
          }
      }
      for (; v < k; v++) {
        const double2 a = to_d2(Vs[(int64_t)v * n + i]);
#pragma unroll
        for (int c = 0; c < NS; c++) {
          const double2 t = cmul(a, hs[c][v]);
          p[c].x -= t.x;
          p[c].y -= t.y;
        }
      }
      double2 q[NS];
#pragma unroll
      for (int c = 0; c < NS; c++) {
        const X o = from_d2<X>(p[c]);
        Z[((int64_t)c * S + s) * n + i] = o;
        q[c] = to_d2(o);
      }
#pragma unroll
      for (int a = 0; a < NS; a++)
#pragma unroll
        for (int b = a; b < NS; b++) {
          g[a][b][0] += q[a].x * q[b].x + q[a].y * q[b].y;
          g[a][b][1] += q[a].x * q[b].y - q[a].y * q[b].x;
        }
    }
  }
  const int wave = threadIdx.x >> 6;
#pragma unroll
  for (int a = 0; a < NS; a++)
#pragma unroll
    for (int b = a; b < NS; b++)
#pragma unroll
      for (int r = 0; r < 2; r++) {
        const double t = wave_sum(g[a][b][r]);
        if ((threadIdx.x & 63) == 0) red[(a * NS + b) * 2 + r][wave] = t;
      }
  __syncthreads();
  if (threadIdx.x < NS * NS * 2) {
    const int ab = threadIdx.x >> 1, a = ab / NS, b = ab % NS, r = threadIdx.x & 1;
    double t = 0.0;
    if (b >= a)
      for (int q = 0; q < kRedThreads / 64; q++) t += red[threadIdx.x][q];
    gram[((int64_t)s * nblk + blockIdx.x) * NS * NS * 2 + threadIdx.x] = t;
  }
}

template <typename X>
void launch_mupdate_s(hipStream_t st, int ns, int S, int64_t n, int k, const X *Vb, int64_t ldv_sys, X *Z,
                      const double *C, int ldc, double *gram, int nblk, const int32_t *active) {
  const dim3 grid(nblk, S);
  if (ns == 2)
    hipLaunchKernelGGL((k_mupdate_s<X, 2>), grid, dim3(kRedThreads), 0, st, n, k, Vb, ldv_sys, Z, C, ldc, gram, nblk,
                       active);
  else if (ns == 3)
    hipLaunchKernelGGL((k_mupdate_s<X, 3>), grid, dim3(kRedThreads), 0, st, n, k, Vb, ldv_sys, Z, C, ldc, gram, nblk,
                       active);
  else if (ns == 4)
    hipLaunchKernelGGL((k_mupdate_s<X, 4>), grid, dim3(kRedThreads), 0, st, n, k, Vb, ldv_sys, Z, C, ldc, gram, nblk,
                       active);
  else
    throw std::runtime_error("mupdate_s: block size must be 2..4");
  HIP_LAUNCH_CHECK();
}
template void launch_mupdate_s<double2>(hipStream_t, int, int, int64_t, int, const double2 *, int64_t, double2 *,
                                        const double *, int, double *, int, const int32_t *);
template void launch_mupdate_s<float2>(hipStream_t, int, int, int64_t, int, const float2 *, int64_t, float2 *,
                                       const double *, int, double *, int, const int32_t *);

// out[s][e] = sum_blk gram[s][blk][e], e < ne (fixed order)
__global__ void k_reduce_small(int nblk, int ne, const double *__restrict__ gram, double *__restrict__ out,
                               const int32_t *__restrict__ active) {
  const int s = blockIdx.x;
  if (active && !active[s]) return;
  const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
  for (int e = wave; e < ne; e += blockDim.x >> 6) {
    double t = 0.0;
    for (int b = l; b < nblk; b += 64) t += gram[((int64_t)s * nblk + b) * ne + e];
    t = wave_sum(t);
    if (l == 0) out[(int64_t)s * ne + e] = t;
  }
}

void launch_reduce_small(hipStream_t st, int S, int nblk, int ne, const double *gram, double *out,
                         const int32_t *active) {
  hipLaunchKernelGGL(k_reduce_small, dim3(S), dim3(256), 0, st, nblk, ne, gram, out, active);
  HIP_LAUNCH_CHECK();
}

// Cholesky factor R (upper, NS x NS, R^H R = G) of the Gram matrix G (upper triangle used); a pivot below 1e-12 of
// its Gram diagonal counts as zero (its column of R is zero, and the matching q the zero vector).
template <int NS>
__device__ __forceinline__ void chol_s(const double *G, double2 R[NS][NS]) {
#pragma unroll
  for (int a = 0; a < NS; a++)
#pragma unroll
    for (int b = 0; b < NS; b++) R[a][b] = make_double2(0.0, 0.0);
#pragma unroll
  for (int a = 0; a < NS; a++) {
    double d = G[(a * NS + a) * 2];
#pragma unroll
    for (int i = 0; i < a; i++) d -= R[i][a].x * R[i][a].x + R[i][a].y * R[i][a].y;
    const double gaa = G[(a * NS + a) * 2];
    const double raa = d > 1e-12 * gaa && d > 0.0 ? sqrt(d) : 0.0;
    R[a][a] = make_double2(raa, 0.0);
#pragma unroll
    for (int b = a + 1; b < NS; b++) {
      double2 t = make_double2(G[(a * NS + b) * 2], G[(a * NS + b) * 2 + 1]);
#pragma unroll
      for (int i = 0; i < a; i++) {
        const double2 u = cconjmul(R[i][a], R[i][b]);
        t.x -= u.x;
        t.y -= u.y;
      }
      R[a][b] = raa > 0.0 ? make_double2(t.x / raa, t.y / raa) : make_double2(0.0, 0.0);
    }
  }
}

// Q = Z R^{-1} (q_c -> Q + c*ldq_sys... of each system), R[s] stored as NS x NS complex
template <typename X, int NS>
__global__ void k_cholqr_s(int64_t n, const double *__restrict__ gram, const X *__restrict__ Z, X *__restrict__ Q,
                           int64_t ldq_sys, double *__restrict__ Rout, const int32_t *__restrict__ active) {
  const int s = blockIdx.y, S = gridDim.y;
  if (active && !active[s]) return;
  double2 R[NS][NS];
  chol_s<NS>(gram + (int64_t)s * NS * NS * 2, R);
  if (blockIdx.x == 0 && threadIdx.x == 0)
    for (int a = 0; a < NS; a++)
      for (int b = 0; b < NS; b++) {
        Rout[(((int64_t)s * NS + a) * NS + b) * 2] = R[a][b].x;
        Rout[(((int64_t)s * NS + a) * NS + b) * 2 + 1] = R[a][b].y;
      }
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double2 q[NS];
#pragma unroll
    for (int c = 0; c < NS; c++) {
      double2 t = to_d2(Z[((int64_t)c * S + s) * n + i]);
#pragma unroll
      for (int a = 0; a < c; a++) {
        const double2 u = cmul(q[a], R[a][c]);
        t.x -= u.x;
        t.y -= u.y;
      }
      const double rc = R[c][c].x;
      q[c] = rc > 0.0 ? make_double2(t.x / rc, t.y / rc) : make_double2(0.0, 0.0);
      Q[s * ldq_sys + (int64_t)c * n + i] = from_d2<X>(q[c]);
    }
  }
}

template <typename X>
void launch_cholqr_s(hipStream_t st, int ns, int S, int64_t n, const double *gram, const X *Z, X *Q, int64_t ldq_sys,
                     double *R, const int32_t *active) {
  int64_t bl = (n + 255) / 256;
  if (bl > 4096) bl = 4096;
  const dim3 grid((unsigned)bl, S);
  if (ns == 2)
    hipLaunchKernelGGL((k_cholqr_s<X, 2>), grid, dim3(256), 0, st, n, gram, Z, Q, ldq_sys, R, active);
  else if (ns == 3)
    hipLaunchKernelGGL((k_cholqr_s<X, 3>), grid, dim3(256), 0, st, n, gram, Z, Q, ldq_sys, R, active);
  else if (ns == 4)
    hipLaunchKernelGGL((k_cholqr_s<X, 4>), grid, dim3(256), 0, st, n, gram, Z, Q, ldq_sys, R, active);
  else
    throw std::runtime_error("cholqr_s: block size must be 2..4");
  HIP_LAUNCH_CHECK();
}
template void launch_cholqr_s<double2>(hipStream_t, int, int, int64_t, const double *, const double2 *, double2 *,
                                       int64_t, double *, const int32_t *);
template void launch_cholqr_s<float2>(hipStream_t, int, int, int64_t, const double *, const float2 *, float2 *,
                                      int64_t, double *, const int32_t *);

// Raw Hessenberg columns j .. j+NS-1 of the s-step relation (z_1 = B v_j, z_{c+1} = B z_c; Z = V C + Q R):
//   column j      = [C_0 ; R_{:,0}]
//   columns j+1.. = [[C_{1:} - Y - C_0 c_j] ; [R_{:,1:} - R_{:,0} c_j]] R_{0:NS-1,0:NS-1}^{-1},
//   c_j = C[j][0:NS-1] (row), Y = Hraw[0..j][0..j-1] C[0..j-1][0:NS-1].
// Both go to Hraw and to the working H; nrm2[t][s] = (R_tt / R_{t-1,t-1})^2 (t >= 1), nrm2[0][s] = R_00^2 are the
// Givens subdiagonals.
template <int NS>
__global__ void k_hcols_s(int m, int j, int S, const double *__restrict__ C, int ldc, const double *__restrict__ Rg,
                          double2 *__restrict__ Hraw, double2 *__restrict__ Hw, double *__restrict__ nrm2,
                          const int32_t *__restrict__ active) {
  const int s = blockIdx.x;
  if (active && !active[s]) return;
  auto cc = [&](int v, int c) {
    const double *p = C + (((int64_t)s * ldc + v) * NS + c) * 2;
    return make_double2(p[0], p[1]);
  };
  double2 R[NS][NS];
#pragma unroll
  for (int a = 0; a < NS; a++)
#pragma unroll
    for (int b = 0; b < NS; b++)
      R[a][b] = make_double2(Rg[(((int64_t)s * NS + a) * NS + b) * 2], Rg[(((int64_t)s * NS + a) * NS + b) * 2 + 1]);
  double2 *Hr = Hraw + (int64_t)s * m * (m + 1);
  double2 *Hs = Hw + (int64_t)s * m * (m + 1);
  double2 cj[NS];
#pragma unroll
  for (int c = 0; c < NS - 1; c++) cj[c] = cc(j, c);
  for (int l = threadIdx.x; l <= j + NS; l += blockDim.x) {
    // column j
    double2 h0 = make_double2(0.0, 0.0);
    if (l <= j)
      h0 = cc(l, 0);
    else
      h0 = R[l - j - 1][0];
    // u = row l of the block before R^{-1}
    double2 u[NS - 1];
    if (l <= j) {
      double2 c0l = cc(l, 0);
#pragma unroll
      for (int c = 0; c < NS - 1; c++) {
        double2 y = make_double2(0.0, 0.0);
        for (int i = (l > 0 ? l - 1 : 0); i < j; i++) {
          const double2 t = cmul(Hr[(int64_t)i * (m + 1) + l], cc(i, c));
          y.x += t.x;
          y.y += t.y;
        }
        const double2 d = cc(l, c + 1), e = cmul(c0l, cj[c]);
        u[c] = make_double2(d.x - y.x - e.x, d.y - y.y - e.y);
      }
    } else {
      const int r = l - j - 1;
#pragma unroll
      for (int c = 0; c < NS - 1; c++) {
        const double2 e = cmul(R[r][0], cj[c]);
        u[c] = make_double2(R[r][c + 1].x - e.x, R[r][c + 1].y - e.y);
      }
    }
    // h = u R_{0:NS-1,0:NS-1}^{-1} (row vector, forward substitution)
    double2 h[NS - 1];
#pragma unroll
    for (int c = 0; c < NS - 1; c++) {
      double2 t = u[c];
#pragma unroll
      for (int a = 0; a < c; a++) {
        const double2 w = cmul(h[a], R[a][c]);
        t.x -= w.x;
        t.y -= w.y;
      }
      const double rd = R[c][c].x;
      h[c] = rd > 0.0 ? make_double2(t.x / rd, t.y / rd) : make_double2(0.0, 0.0);
    }
    if (l <= j + 1) {
      Hr[(int64_t)j * (m + 1) + l] = h0;
      Hs[(int64_t)j * (m + 1) + l] = h0;
    }
#pragma unroll
    for (int c = 0; c < NS - 1; c++) {
      const int col = j + 1 + c;
      if (col < m && l <= col + 1) {
        Hr[(int64_t)col * (m + 1) + l] = h[c];
        Hs[(int64_t)col * (m + 1) + l] = h[c];
      }
    }
  }
  if (threadIdx.x == 0) {
    nrm2[s] = R[0][0].x * R[0][0].x;
#pragma unroll
    for (int t = 1; t < NS; t++) {
      const double q = R[t - 1][t - 1].x > 0.0 ? R[t][t].x / R[t - 1][t - 1].x : 0.0;
      nrm2[(int64_t)t * S + s] = q * q;
    }
  }
}

void launch_hcols_s(hipStream_t st, int ns, int S, int m, int j, const double *C, int ldc, const double *R,
                    double *Hraw, double *Hw, double *nrm2, const int32_t *active) {
  if (ns == 2)
    hipLaunchKernelGGL((k_hcols_s<2>), dim3(S), dim3(256), 0, st, m, j, S, C, ldc, R, (double2 *)Hraw, (double2 *)Hw,
                       nrm2, active);
  else if (ns == 3)
    hipLaunchKernelGGL((k_hcols_s<3>), dim3(S), dim3(256), 0, st, m, j, S, C, ldc, R, (double2 *)Hraw, (double2 *)Hw,
                       nrm2, active);
  else if (ns == 4)
    hipLaunchKernelGGL((k_hcols_s<4>), dim3(S), dim3(256), 0, st, m, j, S, C, ldc, R, (double2 *)Hraw, (double2 *)Hw,
                       nrm2, active);
  else
    throw std::runtime_error("hcols_s: block size must be 2..4");
  HIP_LAUNCH_CHECK();
}

// partial layout for k>=1 complex: [s][blk][kMaxK] double2 ; for norms (k == 0): [s][blk] double.
// Workgroup (s, chunk of 64 entries): lane (v, q) sums blocks q, q+4, ... for entry v; the four sums are then
// added in a fixed order, so the result does not depend on scheduling.
__global__ __launch_bounds__(256) void k_reduce(int nblk, int k, int ldp, const double *__restrict__ partial,
                                                double *__restrict__ out, int ldo, double *__restrict__ Hcol, int ldH,
                                                const int32_t *__restrict__ active) {
  const int s = blockIdx.x;
  if (active && !active[s]) return;
  __shared__ double red[4][64][2];
  const int vl = threadIdx.x & 63, q = threadIdx.x >> 6;
  if (k == 0) {
    double t = 0.0;
    for (int b = threadIdx.x; b < nblk; b += blockDim.x) t += partial[(int64_t)s * nblk + b];
    t = wave_sum(t);
    if (vl == 0) red[q][0][0] = t;
    __syncthreads();
    if (threadIdx.x == 0) out[(int64_t)s * ldo] = red[0][0][0] + red[1][0][0] + red[2][0][0] + red[3][0][0];
    return;
  }
  const int v = blockIdx.y * 64 + vl;
  double tx = 0.0, ty = 0.0;
  if (v < k) {
#pragma unroll 4
    for (int b = q; b < nblk; b += 4) {
      const double *p = partial + (((int64_t)s * nblk + b) * ldp + v) * 2;
      tx += p[0];
      ty += p[1];
    }
  }
  red[q][vl][0] = tx;
  red[q][vl][1] = ty;
  __syncthreads();
  if (q == 0 && v < k) {
    tx = ((red[0][vl][0] + red[1][vl][0]) + red[2][vl][0]) + red[3][vl][0];
    ty = ((red[0][vl][1] + red[1][vl][1]) + red[2][vl][1]) + red[3][vl][1];
    out[((int64_t)s * ldo + v) * 2] = tx;
    out[((int64_t)s * ldo + v) * 2 + 1] = ty;
    if (Hcol) {
      Hcol[((int64_t)s * ldH + v) * 2] += tx;
      Hcol[((int64_t)s * ldH + v) * 2 + 1] += ty;
    }
  }
}

void launch_reduce_ld(hipStream_t st, int S, int nblk, int k, int ldp, const double *partial, double *out, int ldo,
                      double *Hcol, int ldH, const int32_t *active) {
  const int nchunk = k == 0 ? 1 : (k + 63) / 64;
  hipLaunchKernelGGL(k_reduce, dim3(S, nchunk), dim3(256), 0, st, nblk, k, ldp, partial, out, ldo, Hcol, ldH,
                     active);
  HIP_LAUNCH_CHECK();
}

void launch_reduce(hipStream_t st, int S, int nblk, int k, const double *partial, double *out, int ldo,
                   double *Hcol, int ldH, const int32_t *active) {
  launch_reduce_ld(st, S, nblk, k, kMaxK, partial, out, ldo, Hcol, ldH, active);
}

template <typename X>
__global__ void k_scale(int64_t n, const X *__restrict__ w, int64_t ldw_sys, X *__restrict__ v, int64_t ldv_sys,
                        const double *__restrict__ nrm2, const int32_t *__restrict__ active) {
  const int s = blockIdx.y;
  if (active && !active[s]) return;
  const double nr = sqrt(nrm2[s]);
  const double inv = nr > 0.0 ? 1.0 / nr : 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double2 a = to_d2(w[s * ldw_sys + i]);
    v[s * ldv_sys + i] = from_d2<X>(make_double2(a.x * inv, a.y * inv));
  }
}

template <typename X>
void launch_scale(hipStream_t st, int S, int64_t n, const X *w, int64_t ldw_sys, X *v, int64_t ldv_sys,
                  const double *nrm2, const int32_t *active) {
  int64_t bl = (n + 255) / 256;
  if (bl > 4096) bl = 4096;
  hipLaunchKernelGGL((k_scale<X>), dim3((unsigned)bl, S), dim3(256), 0, st, n, w, ldw_sys, v, ldv_sys, nrm2, active);
  HIP_LAUNCH_CHECK();
}
template void launch_scale<double2>(hipStream_t, int, int64_t, const double2 *, int64_t, double2 *, int64_t,
                                    const double *, const int32_t *);
template void launch_scale<float2>(hipStream_t, int, int64_t, const float2 *, int64_t, float2 *, int64_t,
                                   const double *, const int32_t *);

// ------------------------------------------------------------------ block Jacobi
// One wavefront per (block, system): gather nd x nd block of A (masked), invert in LDS with
// partial pivoting (Gauss-Jordan, in place), write owned rows transposed: inv[s][b][r].
template <typename P>
__global__ __launch_bounds__(64) void k_bj_build(int R, int nd, const int32_t *__restrict__ elem_dofs,
                                                 const int32_t *__restrict__ elem_nz, const double2 *__restrict__ a64,
                                                 int64_t nnz, const uint8_t *__restrict__ rowmask, int64_t n,
                                                 const int32_t *__restrict__ owner_elem, P *__restrict__ inv,
                                                 const int32_t *__restrict__ elem_of_block) {
  extern __shared__ double2 lds[];
  double2 *A = lds;                                   // nd*nd
  double2 *colk = lds + nd * nd;                      // nd
  int *perm = (int *)(lds + nd * nd + nd);            // nd
  const int s = blockIdx.y;
  const int e = elem_of_block[blockIdx.x];
  const int g = s / R, r = s % R;
  const int lane = threadIdx.x;
  const int32_t *dofs = elem_dofs + (int64_t)e * nd;
  const double2 *Ag = a64 + (int64_t)g * nnz;
  const uint8_t *mask = rowmask + (int64_t)r * n;
  const int ndd = nd * nd;
  for (int t = lane; t < ndd; t += 64) {
    const int a = t / nd, b = t - a * nd;
    const bool fa = mask[dofs[a]] != 0, fb = mask[dofs[b]] != 0;
    double2 v;
    if (fa && fb)
      v = Ag[elem_nz[(int64_t)e * ndd + t]];
    else
      v = make_double2(a == b ? 1.0 : 0.0, 0.0);
    A[t] = v;
  }
  __syncthreads();
  for (int k = 0; k < nd; k++) {
    // pivot search in column k, rows k..nd-1
    double best = -1.0;
    int bi = k;
    for (int i = k + lane; i < nd; i += 64) {
      const double2 v = A[i * nd + k];
      const double m = v.x * v.x + v.y * v.y;
      if (m > best) {
        best = m;
        bi = i;
      }
    }
    for (int o = 32; o > 0; o >>= 1) {
      const double ob = __shfl_xor(best, o, 64);
      const int oi = __shfl_xor(bi, o, 64);
      if (ob > best || (ob == best && oi < bi)) {
        best = ob;
        bi = oi;
      }
    }
    if (lane == 0) perm[k] = bi;
    if (bi != k)
      for (int c = lane; c < nd; c += 64) {
        const double2 t0 = A[k * nd + c];
        A[k * nd + c] = A[bi * nd + c];
        A[bi * nd + c] = t0;
      }
    __syncthreads();
    const double2 piv = A[k * nd + k];
    const double pm = piv.x * piv.x + piv.y * piv.y;
    const double2 pinv = pm > 0.0 ? make_double2(piv.x / pm, -piv.y / pm) : make_double2(0.0, 0.0);
    __syncthreads();
    for (int c = lane; c < nd; c += 64) A[k * nd + c] = (c == k) ? pinv : cmul(A[k * nd + c], pinv);
    for (int i = lane; i < nd; i += 64) colk[i] = A[i * nd + k];
    __syncthreads();
    for (int t = lane; t < ndd; t += 64) {
      const int i = t / nd, c = t - i * nd;
      if (i == k) continue;
      const double2 f = colk[i];
      if (c == k)
        A[t] = make_double2(-(f.x * pinv.x - f.y * pinv.y), -(f.x * pinv.y + f.y * pinv.x));
      else {
        const double2 u = cmul(f, A[k * nd + c]);
        A[t] = make_double2(A[t].x - u.x, A[t].y - u.y);
      }
    }
    __syncthreads();
  }
  // undo column permutation (reverse order)
  for (int k = nd - 1; k >= 0; k--) {
    const int pk = perm[k];
    if (pk != k)
      for (int i = lane; i < nd; i += 64) {
        const double2 t0 = A[i * nd + k];
        A[i * nd + k] = A[i * nd + pk];
        A[i * nd + pk] = t0;
      }
    __syncthreads();
  }
  // owned rows
  for (int t = lane; t < ndd; t += 64) {
    const int a = t / nd, b = t - a * nd;
    const int dof = dofs[a];
    if (owner_elem[dof] != e) continue;
    // first occurrence of the DOF in the element is its owner slot (a DOF appears once per element)
    inv[((int64_t)s * nd + b) * n + dof] = from_d2<P>(A[a * nd + b]);
  }
}

template <typename P>
void launch_bj_build(hipStream_t st, int S, int R, int nd, const int32_t *elem_dofs, const int32_t *elem_nz,
                     const double *a64, int64_t nnz, const uint8_t *rowmask, int64_t n, const int32_t *owner_elem,
                     P *inv_rows, const int32_t *elem_of_block, int64_t nblocks) {
  size_t lds = (size_t)(nd * nd + nd) * sizeof(double2) + nd * sizeof(int) + 16;
  hipLaunchKernelGGL((k_bj_build<P>), dim3((unsigned)nblocks, S), dim3(64), lds, st, R, nd, elem_dofs, elem_nz,
                     (const double2 *)a64, nnz, rowmask, n, owner_elem, inv_rows, elem_of_block);
  HIP_LAUNCH_CHECK();
}
template void launch_bj_build<double2>(hipStream_t, int, int, int, const int32_t *, const int32_t *, const double *,
                                       int64_t, const uint8_t *, int64_t, const int32_t *, double2 *,
                                       const int32_t *, int64_t);
template void launch_bj_build<float2>(hipStream_t, int, int, int, const int32_t *, const int32_t *, const double *,
                                      int64_t, const uint8_t *, int64_t, const int32_t *, float2 *, const int32_t *,
                                      int64_t);

template <typename P, typename X>
__global__ __launch_bounds__(256) void k_bj_apply(int64_t n, int nd, const int32_t *__restrict__ elem_dofs,
                                                  const int32_t *__restrict__ owner_elem, const P *__restrict__ inv,
                                                  const X *__restrict__ v, int64_t ldv_sys, X *__restrict__ z,
                                                  const int32_t *__restrict__ active, int64_t inv_stride) {
  const int s = blockIdx.y, S = gridDim.y;
  if (active && !active[s]) return;
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int32_t *dofs = elem_dofs + (int64_t)owner_elem[r] * nd;
  const X *vs = v + s * ldv_sys;
  const P *is = inv + (int64_t)s * inv_stride;
  double2 acc = make_double2(0.0, 0.0);
  for (int b = 0; b < nd; b++) {
    const double2 p = to_d2(is[(int64_t)b * n + r]);
    const double2 x = to_d2(vs[dofs[b]]);
    acc.x += p.x * x.x - p.y * x.y;
    acc.y += p.x * x.y + p.y * x.x;
  }
  z[r * S + s] = from_d2<X>(acc);
}

template <typename P, typename X>
void launch_bj_apply(hipStream_t st, int S, int64_t n, int nd, const int32_t *elem_dofs, const int32_t *owner_elem,
                     const P *inv_rows, const X *v, int64_t ldv_sys, X *z, const int32_t *active, int64_t inv_stride) {
  hipLaunchKernelGGL((k_bj_apply<P, X>), dim3((unsigned)((n + 255) / 256), S), dim3(256), 0, st, n, nd, elem_dofs,
                     owner_elem, inv_rows, v, ldv_sys, z, active, inv_stride);
  HIP_LAUNCH_CHECK();
}
template void launch_bj_apply<double2, double2>(hipStream_t, int, int64_t, int, const int32_t *, const int32_t *,
                                                const double2 *, const double2 *, int64_t, double2 *,
                                                const int32_t *, int64_t);
template void launch_bj_apply<float2, float2>(hipStream_t, int, int64_t, int, const int32_t *, const int32_t *,
                                              const float2 *, const float2 *, int64_t, float2 *, const int32_t *,
                                              int64_t);

// ------------------------------------------------------------------ vertex-patch additive Schwarz
// Storage per preconditioner group q (pg consecutive frequency groups): [shared region: patches with no
// port DOF, one inverse for all excitations] [R excitation regions: patches touching a port DOF].
// Element offset of patch p, group q, excitation r:
//   q*group_stride + (port[p] ? shared_size + r*exc_size : 0) + val_off[p].
// FP64 storage: the full n_p x n_p inverse, row-major.  FP16 storage: the symmetrised inverse as two
// row-major planes of halves, real then imaginary, each n_p x npk with npk = n_p rounded up to a multiple
// of 8 (zero columns n_p..npk-1, so every row starts 16-byte aligned); block size n_p*npk in __half2
// units.  FP16 blocks carry one scale per (group, excitation, patch): value = half * scale.
__device__ __forceinline__ int64_t patch_offset(const PatchLayout &L, int p, int q, int r) {
  return (int64_t)q * L.group_stride + (L.port[p] ? L.shared_size + (int64_t)r * L.exc_size : 0) + L.val_off[p];
}
__device__ __forceinline__ int pad8(int n) { return (n + 7) & ~7; }

// Gauss-Jordan with partial pivoting in LDS, in FP64 (n_p <= 100: 160 KB).
// FP16 storage keeps a block as four planes of n_p x roundup8(n_p) halves, [Re hi | Im hi | Re lo | Im lo], of the
// symmetrised inverse scaled so its largest |Re|,|Im| is 1: hi = f16(a), lo = f16(2^11 (a - hi)).  The hi planes
// alone carry 11 bits, enough while the block's condition estimate kappa = |A_p|_F |A_p^-1|_F / n_p stays small;
// above the threshold (low k0 h: the near-kernel gradient part dominates the inverse by orders of magnitude and
// hi-only rounding destroys the rest) the block is marked by a negative scale and applied with both planes
// (about 22 bits).  Blocks below it leave the lo planes unwritten and unread.
template <typename T>
struct Cx;
template <>
struct Cx<float> {
  using V = float2;
};
template <>
struct Cx<double> {
  using V = double2;
};

// One workgroup of kBuildThreads threads per block: the FP64 block fills the LDS of a CU, so the workgroup
// itself must carry the CU's parallelism (16 wavefronts, four per SIMD).
constexpr int kBuildThreads = 1024;
constexpr int kBuildWaves = kBuildThreads / 64;

template <typename P>
__global__ __launch_bounds__(kBuildThreads) void k_patch_build(PatchLayout L, int G, int R, const double2 *__restrict__ a64,
                                                     int64_t nnz, const uint8_t *__restrict__ rowmask, int64_t n,
                                                     P *__restrict__ inv, float *__restrict__ scale,
                                                     float split_kappa) {
  constexpr bool kPacked = std::is_same<P, __half2>::value;
  using T = double;
  using C2 = typename Cx<T>::V;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  C2 *lds = (C2 *)lds_raw;
  const int p = blockIdx.x;
  const int q = blockIdx.y / R, r = blockIdx.y % R;
  if (!L.port[p] && r != 0) return;  // shared inverse: built once per group
  const int g0 = q * L.pg, ng = min(L.pg, G - g0);
  const double wg = 1.0 / ng;
  const int np = L.idx_off[p + 1] - L.idx_off[p];
  C2 *A = lds;                 // np*np
  C2 *colk = lds + np * np;    // np
  __shared__ int perm[128];
  __shared__ C2 pinv_s;
  __shared__ double red[kBuildWaves];
  const int32_t *I = L.idx + L.idx_off[p];
  const int32_t *nz = L.pnz + L.nz_off[p];
  const uint8_t *mask = rowmask + (int64_t)r * n;
  const int tid = threadIdx.x;
  const int nn = np * np;
  auto mk = [](T x, T y) {
    C2 o;
    o.x = x;
    o.y = y;
    return o;
  };
  auto mul = [&](C2 a, C2 b) { return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); };
  // block of the mean matrix of the group's frequencies (the frequencies of a group are adjacent)
  for (int t = tid; t < nn; t += blockDim.x) {
    const int a = t / np, b = t - a * np;
    const bool fa = mask[I[a]] != 0, fb = mask[I[b]] != 0;
    double vx = a == b ? 1.0 : 0.0, vy = 0.0;
    if (fa && fb) {
      const int32_t qz = nz[t];
      vx = 0.0;
      if (qz >= 0)
        for (int g = g0; g < g0 + ng; g++) {
          const double2 e = a64[(int64_t)g * nnz + qz];
          vx += e.x;
          vy += e.y;
        }
      vx *= wg;
      vy *= wg;
    }
    A[t] = mk((T)vx, (T)vy);
  }
  __syncthreads();
  double a2 = 0.0;  // |A_p|_F^2 (FP16 storage: condition estimate)
  if constexpr (kPacked) {
    for (int t = tid; t < nn; t += blockDim.x) a2 += A[t].x * A[t].x + A[t].y * A[t].y;
    a2 = wave_sum(a2);
    if ((tid & 63) == 0) red[tid >> 6] = a2;
    __syncthreads();
    a2 = 0.0;
    for (int w = 0; w < kBuildWaves; w++) a2 += red[w];
    __syncthreads();
  }
  for (int k = 0; k < np; k++) {
    if (tid < 64) {
      T best = (T)-1;
      int bi = k;
      for (int i = k + tid; i < np; i += 64) {
        const C2 v = A[i * np + k];
        const T m = v.x * v.x + v.y * v.y;
        if (m > best) {
          best = m;
          bi = i;
        }
      }
      for (int o = 32; o > 0; o >>= 1) {
        const T ob = __shfl_xor(best, o, 64);
        const int oi = __shfl_xor(bi, o, 64);
        if (ob > best || (ob == best && oi < bi)) {
          best = ob;
          bi = oi;
        }
      }
      if (tid == 0) perm[k] = bi;
    }
    __syncthreads();
    const int bi = perm[k];
    if (bi != k)
      for (int c = tid; c < np; c += blockDim.x) {
        const C2 t0 = A[k * np + c];
        A[k * np + c] = A[bi * np + c];
        A[bi * np + c] = t0;
      }
    __syncthreads();
    if (tid == 0) {
      const C2 piv = A[k * np + k];
      const T pm = piv.x * piv.x + piv.y * piv.y;
      pinv_s = pm > (T)0 ? mk(piv.x / pm, -piv.y / pm) : mk((T)0, (T)0);
    }
    __syncthreads();
    const C2 pinv = pinv_s;
    for (int c = tid; c < np; c += blockDim.x) A[k * np + c] = (c == k) ? pinv : mul(A[k * np + c], pinv);
    for (int i = tid; i < np; i += blockDim.x) colk[i] = A[i * np + k];
    __syncthreads();
    // rows over wavefronts, columns over lanes (np <= 100: two columns per lane)
    for (int i = tid >> 6; i < np; i += blockDim.x >> 6) {
      if (i == k) continue;
      const C2 f = colk[i];
      for (int c = tid & 63; c < np; c += 64) {
        const int t = i * np + c;
        if (c == k) {
          A[t] = mk(-(f.x * pinv.x - f.y * pinv.y), -(f.x * pinv.y + f.y * pinv.x));
        } else {
          const C2 u = mul(f, A[k * np + c]);
          A[t] = mk(A[t].x - u.x, A[t].y - u.y);
        }
      }
    }
    __syncthreads();
  }
  for (int k = np - 1; k >= 0; k--) {
    const int pk = perm[k];
    if (pk != k)
      for (int i = tid; i < np; i += blockDim.x) {
        const C2 t0 = A[i * np + k];
        A[i * np + k] = A[i * np + pk];
        A[i * np + pk] = t0;
      }
    __syncthreads();
  }
  P *out = inv + patch_offset(L, p, q, r);
  if constexpr (kPacked) {
    // block scale: the largest |Re|,|Im| maps to 1 in the half-precision store; |A_p^-1|_F for the condition
    // estimate
    double m = 0.0, i2 = 0.0;
    for (int t = tid; t < nn; t += blockDim.x) {
      m = fmax(m, fmax(fabs(A[t].x), fabs(A[t].y)));
      i2 += A[t].x * A[t].x + A[t].y * A[t].y;
    }
    m = wave_sum_max(m);
    i2 = wave_sum(i2);
    __shared__ double red2[kBuildWaves];
    if ((tid & 63) == 0) {
      red[tid >> 6] = m;
      red2[tid >> 6] = i2;
    }
    __syncthreads();
    m = 0.0;
    i2 = 0.0;
    for (int w = 0; w < kBuildWaves; w++) {
      m = fmax(m, red[w]);
      i2 += red2[w];
    }
    const double sc = m > 0.0 ? m : 1.0;
    const bool split = sqrt(a2 * i2) / np > (double)split_kappa;
    if (tid == 0) scale[((int64_t)q * R + (L.port[p] ? r : 0)) * L.npatch + p] = split ? -(float)sc : (float)sc;
    const double h = 0.5 / sc;
    const int npk = pad8(np);
    const int64_t pl = (int64_t)np * npk;
    __half *Pr = (__half *)out, *Pi = Pr + pl, *Prl = Pr + 2 * pl, *Pil = Pr + 3 * pl;
    for (int t = tid; t < np * npk; t += blockDim.x) {
      const int i = t / npk, j = t - i * npk;
      double re = 0.0, im = 0.0;
      if (j < np) {
        const C2 u = A[i * np + j], w = A[j * np + i];
        re = (u.x + w.x) * h;
        im = (u.y + w.y) * h;
      }
      const __half hr = __float2half((float)re), hi = __float2half((float)im);
      Pr[t] = hr;
      Pi[t] = hi;
      if (split) {
        Prl[t] = __float2half((float)((re - (double)__half2float(hr)) * 2048.0));
        Pil[t] = __float2half((float)((im - (double)__half2float(hi)) * 2048.0));
      }
    }
  } else {
    for (int t = tid; t < nn; t += blockDim.x) out[t] = make_double2((double)A[t].x, (double)A[t].y);
  }
}

// out[q][e] = mean of a64[g][e] over the frequencies g of group q (pg per group, the last may hold fewer):
// the matrix a shared set of patch inverses is built from, read once as a stream
__global__ __launch_bounds__(256) void k_group_mean(int G, int pg, int64_t nnz, const double2 *__restrict__ a64,
                                                    double2 *__restrict__ out) {
  const int q = blockIdx.y;
  const int g0 = q * pg, ng = min(pg, G - g0);
  const double w = 1.0 / ng;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nnz; e += (int64_t)gridDim.x * blockDim.x) {
    double2 acc = make_double2(0.0, 0.0);
    for (int g = g0; g < g0 + ng; g++) {
      const double2 a = a64[(int64_t)g * nnz + e];
      acc.x += a.x;
      acc.y += a.y;
    }
    out[(int64_t)q * nnz + e] = make_double2(acc.x * w, acc.y * w);
  }
}

void launch_group_mean(hipStream_t st, int G, int pg, int64_t nnz, const double *a64, double *out) {
  const int nq = (G + pg - 1) / pg;
  const unsigned bx = (unsigned)std::min<int64_t>(4096, (nnz + 255) / 256);
  hipLaunchKernelGGL(k_group_mean, dim3(bx, nq), dim3(256), 0, st, G, pg, nnz, (const double2 *)a64, (double2 *)out);
  HIP_LAUNCH_CHECK();
}

template <typename P>
void launch_patch_build(hipStream_t st, const PatchLayout &L, int G, int R, const double *a64, int64_t nnz,
                        const uint8_t *rowmask, int64_t n, P *inv, float *scale) {
  if (L.nmax > 100) throw std::runtime_error("vertex patch larger than 100 DOFs");
  if (L.pg < 1) throw std::runtime_error("patch build: frequencies per preconditioner must be >= 1");
  const size_t lds = (size_t)(L.nmax * L.nmax + L.nmax) * 16;
  static size_t attr_bytes = 0;
  if (lds > attr_bytes) {
    hipError_t e = hipFuncSetAttribute((const void *)k_patch_build<P>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)lds);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      throw std::runtime_error(std::string("patch build: LDS request failed: ") + hipGetErrorString(e));
    }
    attr_bytes = lds;
  }
  const int nq = (G + L.pg - 1) / L.pg;
  hipLaunchKernelGGL((k_patch_build<P>), dim3((unsigned)L.npatch, nq * R), dim3(kBuildThreads), lds, st, L, G, R,
                     (const double2 *)a64, nnz, rowmask, n, inv, scale, patch_split_kappa());
  HIP_LAUNCH_CHECK();
}
template void launch_patch_build<double2>(hipStream_t, const PatchLayout &, int, int, const double *, int64_t,
                                          const uint8_t *, int64_t, double2 *, float *);
template void launch_patch_build<__half2>(hipStream_t, const PatchLayout &, int, int, const double *, int64_t,
                                          const uint8_t *, int64_t, __half2 *, float *);

// Apply in two deterministic passes.
// 1) one workgroup per (patch, preconditioner group): y_p = inv_p v_p for the group's ns = pg*R systems,
//    v_p staged in LDS.  The inverse of a complex-symmetric block is complex symmetric, so thread t forms
//    row t as column t, reading A[j][t] (coalesced across t).  A patch touching a port DOF has one
//    inverse per excitation r, applied to the systems with s % R == r.  FP16 storage computes in FP32.
// 2) one thread per DOF: z[d] = sum over the (patch, row) pairs of d of y.
template <typename P, typename X, int NSM>
__global__ __launch_bounds__(128) void k_patch_local(PatchLayout L, int G, int R, const P *__restrict__ inv,
                                                     const float *__restrict__ scale, const X *__restrict__ v,
                                                     int64_t ldv_sys, X *__restrict__ y, int64_t ldy_sys,
                                                     const int32_t *__restrict__ active, int group_fast) {
  // FP32 arithmetic for FP32 vectors; FP64 arithmetic (a linear operator to FP64 rounding) for FP64 vectors
  using T = typename std::conditional<std::is_same<X, float2>::value, float, double>::type;
  // group_fast: the groups of a patch are consecutive workgroups (grid (groups, patches)), so groups that share
  // one inverse (group stride 0) read it from L2 after the first
  const int p = group_fast ? blockIdx.y : blockIdx.x, q = group_fast ? blockIdx.x : blockIdx.y;
  const int g0 = q * L.pg;
  const int ns = min(L.pg, G - g0) * R;
  const int64_t s0 = (int64_t)g0 * R;
  bool act[NSM];
  bool any = false;
#pragma unroll
  for (int k = 0; k < NSM; k++) {
    act[k] = k < ns && (active == nullptr || active[s0 + k]);
    any |= act[k];
  }
  if (!any) return;
  __shared__ X vs[NSM][128];
  const int b0 = L.idx_off[p], np = L.idx_off[p + 1] - b0;
  const int t = threadIdx.x;
  if (t < np) {
    const int32_t c = L.idx[b0 + t];
#pragma unroll
    for (int k = 0; k < NSM; k++)
      if (k < ns) vs[k][t] = v[(s0 + k) * ldv_sys + c];
  }
  __syncthreads();
  if (t >= np) return;
  const int64_t yo = L.yoff[p] + t;
  const int nreg = L.port[p] ? R : 1;
  const int npk = pad8(np);
  for (int ri = 0; ri < nreg; ri++) {
    const P *A = inv + patch_offset(L, p, q, ri);
    const float scv = scale ? scale[((int64_t)q * R + ri) * L.npatch + p] : 1.f;
    const bool lo = scv < 0.f;  // FP16: the block carries its lo planes
    const bool real = std::is_same<P, __half2>::value && L.real_inv != 0;  // uniform
    const T sc = (T)fabsf(scv);
    T ax[NSM], ay[NSM];
#pragma unroll
    for (int k = 0; k < NSM; k++) ax[k] = ay[k] = (T)0;
    auto entry = [&](int j, T &a_x, T &a_y) {
      if constexpr (std::is_same<P, __half2>::value) {
        const __half *Pr = (const __half *)A;
        const int64_t pl = (int64_t)np * npk, e = (int64_t)j * npk + t;
        a_x = __half2float(Pr[e]);
        a_y = real ? (T)0 : (T)__half2float(Pr[pl + e]);
        if (lo) {
          a_x += __half2float(Pr[2 * pl + e]) * (T)(1.0 / 2048.0);
          if (!real) a_y += __half2float(Pr[3 * pl + e]) * (T)(1.0 / 2048.0);
        }
      } else {
        const double2 a = A[(int64_t)j * np + t];
        a_x = a.x;
        a_y = a.y;
      }
    };
    auto fma_row = [&](int j, T a_x, T a_y) {
#pragma unroll
      for (int k = 0; k < NSM; k++) {
        if (k < ns) {
          const X xv = vs[k][j];
          ax[k] += a_x * (T)xv.x - a_y * (T)xv.y;
          ay[k] += a_x * (T)xv.y + a_y * (T)xv.x;
        }
      }
    };
    auto fma_row_real = [&](int j, T a_x) {
#pragma unroll
      for (int k = 0; k < NSM; k++) {
        if (k < ns) {
          const X xv = vs[k][j];
          ax[k] += a_x * (T)xv.x;
          ay[k] += a_x * (T)xv.y;
        }
      }
    };
    // four rows of the inverse are loaded before any is used (four loads in flight per lane)
    int j = 0;
    for (; j + 4 <= np; j += 4) {
      T a_x[4], a_y[4];
#pragma unroll
      for (int u = 0; u < 4; u++) entry(j + u, a_x[u], a_y[u]);
      if (real) {
#pragma unroll
        for (int u = 0; u < 4; u++) fma_row_real(j + u, a_x[u]);
      } else {
#pragma unroll
        for (int u = 0; u < 4; u++) fma_row(j + u, a_x[u], a_y[u]);
      }
    }
    for (; j < np; j++) {
      T a_x, a_y;
      entry(j, a_x, a_y);
      fma_row(j, a_x, a_y);
    }
#pragma unroll
    for (int k = 0; k < NSM; k++) {
      const bool mine = nreg == 1 || (k % R) == ri;
      if (act[k] && mine) {
        X o;
        o.x = sc * ax[k];
        o.y = sc * ay[k];
        y[yo * ldy_sys + s0 + k] = o;
      }
    }
  }
}

// FP16 storage on MFMA: one workgroup per (patch, preconditioner group), one wavefront per 8 systems of the
// group (up to four wavefronts read the same inverse on one CU, so the block comes from HBM once).  The
// complex product is two real products, [Yr | Yi] = Ar [Vr | Vi] + Ai [-Vi | Vr], an
// (n_p x n_p) x (n_p x 2ns) GEMM on v_mfma_f32_16x16x32_f16 (FP32 accumulation): column j < ns is Re of
// system j, ns <= j < 2ns is Im of system j - ns (ns <= 8 systems of the wavefront).  Each system's patch
// vector is scaled to max |Re|,|Im| = 1 and split into two halves, hi = f16(v), lo = f16(v - hi), so the
// operand keeps about 22 bits; the inverse is exact in its FP16 store.  Operand layouts (16x16x32, 8 values
// per lane): A[i][k] in lane i%16 + 16*(k%32/8), B[k][j] in lane j + 16*(k%32/8), D[i][j] in register i%4 of
// lane j + 16*(i/4).
typedef _Float16 h8_t __attribute__((ext_vector_type(8)));
typedef float f4v_t __attribute__((ext_vector_type(4)));
constexpr int kPatchMfmaSys = 32;  // systems per (patch, group) on the MFMA path: up to four wavefronts

template <typename X>
__global__ __launch_bounds__(256) void k_patch_local_mfma(PatchLayout L, int G, int R, const __half2 *__restrict__ inv,
                                                          const float *__restrict__ scale, const X *__restrict__ v,
                                                          int64_t ldv_sys, X *__restrict__ y, int64_t ldy_sys,
                                                          const int32_t *__restrict__ active) {
  const int p = blockIdx.x, q = blockIdx.y;
  const int g0 = q * L.pg;
  const int ns = min(L.pg, G - g0) * R;  // <= 8 * (blockDim.x / 64)
  const int64_t s0 = (int64_t)g0 * R;
  bool any = false;
  for (int k = 0; k < ns; k++) any |= active == nullptr || active[s0 + k] != 0;
  if (!any) return;  // uniform over the workgroup
  __shared__ float2 vs[kPatchMfmaSys][128];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int b0 = L.idx_off[p], np = L.idx_off[p + 1] - b0, npk = pad8(np);
  {
    // the group's patch vectors into LDS, entry e = 128 sy + t (zero past n_p); blockDim = 64 ceil(ns / 8),
    // so 16 entries per thread cover the group.  Indices are clamped so every load is issued unconditionally.
    const int ne = ns * 128;
    int32_t col[16];
#pragma unroll
    for (int m = 0; m < 16; m++) col[m] = L.idx[b0 + min((int)((threadIdx.x + m * blockDim.x) & 127), np - 1)];
    X val[16];
#pragma unroll
    for (int m = 0; m < 16; m++) {
      const int e = min((int)(threadIdx.x + m * blockDim.x), ne - 1);
      val[m] = v[(s0 + (e >> 7)) * ldv_sys + col[m]];
    }
#pragma unroll
    for (int m = 0; m < 16; m++) {
      const int e = threadIdx.x + m * blockDim.x, t = e & 127;
      if (e < ne) vs[e >> 7][t] = t < np ? make_float2((float)val[m].x, (float)val[m].y) : make_float2(0.f, 0.f);
    }
  }
  __syncthreads();
  // this wavefront's systems: [8 wv, 8 wv + nw) of the group
  const int nw = min(8, ns - 8 * wv);
  const int j = lane & 15, kq = lane >> 4;
  const int sl = j < 2 * nw ? (j < nw ? j : j - nw) : -1;
  const int sy = sl >= 0 ? 8 * wv + sl : -1;
  const int comp = j < nw ? 0 : 1;
  // max |Re|,|Im| of the system's patch vector: the four lanes j, j+16, j+32, j+48 hold its 128 entries
  float ms = 0.f;
  if (sy >= 0)
#pragma unroll
    for (int kb = 0; kb < 4; kb++)
#pragma unroll
      for (int e = 0; e < 8; e++) {
        const float2 x = vs[sy][kb * 32 + kq * 8 + e];
        ms = fmaxf(ms, fmaxf(fabsf(x.x), fabsf(x.y)));
      }
  ms = fmaxf(ms, __shfl_xor(ms, 16, 64));
  ms = fmaxf(ms, __shfl_xor(ms, 32, 64));
  // below the smallest normal FP32 the reciprocal would overflow (0 * inf = NaN); such a patch vector is
  // negligible against the unit-norm Krylov vector and contributes zero
  if (ms < 1.17549435e-38f) ms = 0.f;
  const float ims = ms > 0.f ? 1.f / ms : 0.f;
  const int nkb = (npk + 31) >> 5;  // <= 4 for n_p <= 100
  h8_t b1h[4], b1l[4], b2h[4], b2l[4];
#pragma unroll
  for (int kb = 0; kb < 4; kb++) {
#pragma unroll
    for (int e = 0; e < 8; e++) {
      const int k = kb * 32 + kq * 8 + e;
      float2 x = make_float2(0.f, 0.f);
      if (sy >= 0) x = vs[sy][k];
      x.x *= ims;
      x.y *= ims;
      const float c1 = comp == 0 ? x.x : x.y;
      const float c2 = comp == 0 ? -x.y : x.x;
      const _Float16 h1 = (_Float16)c1, h2 = (_Float16)c2;
      b1h[kb][e] = h1;
      b1l[kb][e] = (_Float16)(c1 - (float)h1);
      b2h[kb][e] = h2;
      b2l[kb][e] = (_Float16)(c2 - (float)h2);
    }
  }
  if (nw <= 0) return;  // no barrier follows
  const int nreg = L.port[p] ? R : 1;
  for (int ri = 0; ri < nreg; ri++) {
    const _Float16 *Ar = (const _Float16 *)(inv + patch_offset(L, p, q, ri));
    const _Float16 *Ai = Ar + (int64_t)np * npk;
    const float scv = scale[((int64_t)q * R + ri) * L.npatch + p];
    const bool lo = scv < 0.f;  // the block carries lo planes (uniform over the workgroup)
    const float sc = fabsf(scv);
    const int64_t s = s0 + sy;
    const bool mine = sy >= 0 && (nreg == 1 || sy % R == ri) && (active == nullptr || active[s] != 0);
    for (int mt = 0; mt < (np + 15) >> 4; mt++) {
      // rows past n_p read row n_p - 1 (their outputs are not stored); columns past the padded width are
      // never read (kb < nkb), and padded columns are zero in the store
      const int i = min(mt * 16 + j, np - 1);
      // the tile's hi and (on marked blocks) lo planes are loaded before any is used: one HBM round trip per
      // row tile
      const int64_t pl = (int64_t)np * npk;
      h8_t ar[4], ai[4], arl[4], ail[4];
#pragma unroll
      for (int kb = 0; kb < 4; kb++)
        if (kb < nkb) {
          const int64_t o = (int64_t)i * npk + min(kb * 32 + kq * 8, npk - 8);  // k0 >= npk meets zero B rows
          ar[kb] = *(const h8_t *)(Ar + o);
          ai[kb] = *(const h8_t *)(Ai + o);
          if (lo) {
            arl[kb] = *(const h8_t *)(Ar + 2 * pl + o);
            ail[kb] = *(const h8_t *)(Ar + 3 * pl + o);
          }
        }
      f4v_t acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kb = 0; kb < 4; kb++) {
        if (kb < nkb) {
          acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(ar[kb], b1h[kb], acc, 0, 0, 0);
          acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(ar[kb], b1l[kb], acc, 0, 0, 0);
          acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(ai[kb], b2h[kb], acc, 0, 0, 0);
          acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(ai[kb], b2l[kb], acc, 0, 0, 0);
        }
      }
      if (lo) {
        // lo planes (2^11 (a - hi)) times the hi halves of the vectors; the lo x lo term is below 2^-22
        f4v_t acc2 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kb = 0; kb < 4; kb++)
          if (kb < nkb) {
            acc2 = __builtin_amdgcn_mfma_f32_16x16x32_f16(arl[kb], b1h[kb], acc2, 0, 0, 0);
            acc2 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ail[kb], b2h[kb], acc2, 0, 0, 0);
          }
#pragma unroll
        for (int r = 0; r < 4; r++) acc[r] += acc2[r] * (1.f / 2048.f);
      }
      if (mine) {
        const float f = sc * ms;
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int row = mt * 16 + kq * 4 + r;
          if (row < np) {
            float *o = (float *)(y + (L.yoff[p] + row) * ldy_sys + s);
            o[comp] = acc[r] * f;
          }
        }
      }
    }
  }
}

template <typename X>
__global__ __launch_bounds__(256) void k_patch_gather(PatchLayout L, int64_t n, int S, const X *__restrict__ y,
                                                      X *__restrict__ z, const int32_t *__restrict__ active) {
  const int64_t tot = n * S;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (int64_t)gridDim.x * blockDim.x) {
    const int s = (int)(e % S);
    if (active && !active[s]) continue;
    const int64_t d = e / S;
    double2 acc = make_double2(0.0, 0.0);
    for (int k = L.d2p_ptr[d]; k < L.d2p_ptr[d + 1]; k++) {
      const X t = y[(int64_t)L.d2p_y[k] * S + s];
      acc.x += t.x;
      acc.y += t.y;
    }
    z[e] = from_d2<X>(acc);
  }
}

template <typename P, typename X>
void launch_patch_apply(hipStream_t st, const PatchLayout &L, int ngroups, int R, int64_t n, const P *inv,
                        const float *scale, const X *v, int64_t ldv_sys, X *z, const int32_t *active, void *ybuf) {
  if (L.nmax > 100) throw std::runtime_error("patch apply: patch larger than 100 DOFs");
  const int nsm = L.pg * R;
  const int nq = (ngroups + L.pg - 1) / L.pg;
  dim3 grid((unsigned)L.npatch, nq);
  X *y = (X *)ybuf;
  const int S = ngroups * R;
  const unsigned gblocks = (unsigned)std::min<int64_t>(65536, (n * S + 255) / 256);
  if constexpr (std::is_same<P, __half2>::value && std::is_same<X, float2>::value) {
    if (patch_mfma_enabled()) {
      if (nsm > kPatchMfmaSys)
        throw std::runtime_error("patch apply: frequencies per preconditioner x excitations must be <= 32");
      const unsigned threads = 64u * (unsigned)((nsm + 7) / 8);
      hipLaunchKernelGGL((k_patch_local_mfma<X>), grid, dim3(threads), 0, st, L, ngroups, R, inv, scale, v, ldv_sys,
                         y, (int64_t)S, active);
      HIP_LAUNCH_CHECK();
      hipLaunchKernelGGL((k_patch_gather<X>), dim3(gblocks), dim3(256), 0, st, L, n, S, y, z, active);
      HIP_LAUNCH_CHECK();
      return;
    }
  }
  if (nsm > 8) throw std::runtime_error("patch apply: frequencies per preconditioner x excitations must be <= 8");
  const int gf = L.group_stride == 0 && nq > 1 && L.npatch <= 65535;
  const dim3 gl = gf ? dim3((unsigned)nq, (unsigned)L.npatch) : grid;
  if (nsm <= 2)
    hipLaunchKernelGGL((k_patch_local<P, X, 2>), gl, dim3(128), 0, st, L, ngroups, R, inv, scale, v, ldv_sys, y,
                       (int64_t)S, active, gf);
  else if (nsm <= 4)
    hipLaunchKernelGGL((k_patch_local<P, X, 4>), gl, dim3(128), 0, st, L, ngroups, R, inv, scale, v, ldv_sys, y,
                       (int64_t)S, active, gf);
  else if (nsm <= 8)
    hipLaunchKernelGGL((k_patch_local<P, X, 8>), gl, dim3(128), 0, st, L, ngroups, R, inv, scale, v, ldv_sys, y,
                       (int64_t)S, active, gf);
  else
    throw std::runtime_error("patch apply: frequencies per preconditioner x excitations must be <= 8");
  HIP_LAUNCH_CHECK();
  hipLaunchKernelGGL((k_patch_gather<X>), dim3(gblocks), dim3(256), 0, st, L, n, S, y, z, active);
  HIP_LAUNCH_CHECK();
}
template void launch_patch_apply<double2, double2>(hipStream_t, const PatchLayout &, int, int, int64_t,
                                                   const double2 *, const float *, const double2 *, int64_t,
                                                   double2 *, const int32_t *, void *);
template void launch_patch_apply<__half2, float2>(hipStream_t, const PatchLayout &, int, int, int64_t,
                                                  const __half2 *, const float *, const float2 *, int64_t, float2 *,
                                                  const int32_t *, void *);
template void launch_patch_apply<__half2, double2>(hipStream_t, const PatchLayout &, int, int, int64_t,
                                                   const __half2 *, const float *, const double2 *, int64_t,
                                                   double2 *, const int32_t *, void *);

// ------------------------------------------------------------------ coarse space
template <typename X>
__global__ __launch_bounds__(256) void k_coarse_restrict(int nc, const int32_t *__restrict__ col_ptr,
                                                         const int32_t *__restrict__ col_row,
                                                         const double *__restrict__ col_val, const X *__restrict__ v,
                                                         int64_t ldv_sys, const X *__restrict__ t, int64_t ldt_sys,
                                                         double2 *__restrict__ rc, const int32_t *__restrict__ active) {
  const int a = blockIdx.x, s = blockIdx.y;
  if (active && !active[s]) return;
  __shared__ double red[2][4];
  const X *vs = v + s * ldv_sys;
  double sx = 0.0, sy = 0.0;
  const X *ts = t ? t + s * ldt_sys : nullptr;
  for (int k = col_ptr[a] + threadIdx.x; k < col_ptr[a + 1]; k += blockDim.x) {
    double2 x = to_d2(vs[col_row[k]]);
    if (ts) {
      const double2 y = to_d2(ts[col_row[k]]);
      x.x -= y.x;
      x.y -= y.y;
    }
    const double c = col_val[k];
    sx += c * x.x;
    sy += c * x.y;
  }
  sx = wave_sum(sx);
  sy = wave_sum(sy);
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = sx;
    red[1][threadIdx.x >> 6] = sy;
  }
  __syncthreads();
  if (threadIdx.x == 0)
    rc[(int64_t)s * nc + a] = make_double2(red[0][0] + red[0][1] + red[0][2] + red[0][3],
                                           red[1][0] + red[1][1] + red[1][2] + red[1][3]);
}

template <typename X>
void launch_coarse_restrict(hipStream_t st, int S, int nc, const int32_t *col_ptr, const int32_t *col_row,
                            const double *col_val, const X *v, int64_t ldv_sys, const X *t, int64_t ldt_sys,
                            double *rc, const int32_t *active) {
  hipLaunchKernelGGL((k_coarse_restrict<X>), dim3(nc, S), dim3(256), 0, st, nc, col_ptr, col_row, col_val, v, ldv_sys,
                     t, ldt_sys, (double2 *)rc, active);
  HIP_LAUNCH_CHECK();
}
template void launch_coarse_restrict<double2>(hipStream_t, int, int, const int32_t *, const int32_t *, const double *,
                                              const double2 *, int64_t, const double2 *, int64_t, double *,
                                              const int32_t *);
template void launch_coarse_restrict<float2>(hipStream_t, int, int, const int32_t *, const int32_t *, const double *,
                                             const float2 *, int64_t, const float2 *, int64_t, double *,
                                             const int32_t *);

__global__ void k_coarse_solve(int R, int nc, const double2 *__restrict__ Ainv, const double2 *__restrict__ rc,
                               double2 *__restrict__ yc, const int32_t *__restrict__ active) {
  const int s = blockIdx.x;
  if (active && !active[s]) return;
  const double2 *A = Ainv + (int64_t)(s / R) * nc * nc;
  for (int i = threadIdx.x; i < nc; i += blockDim.x) {
    double2 acc = make_double2(0.0, 0.0);
    for (int j = 0; j < nc; j++) {
      const double2 u = cmul(A[(int64_t)i * nc + j], rc[(int64_t)s * nc + j]);
      acc.x += u.x;
      acc.y += u.y;
    }
    yc[(int64_t)s * nc + i] = acc;
  }
}

void launch_coarse_solve(hipStream_t st, int S, int R, int nc, const double *Ainv, const double *rc, double *yc,
                         const int32_t *active) {
  hipLaunchKernelGGL(k_coarse_solve, dim3(S), dim3(128), 0, st, R, nc, (const double2 *)Ainv, (const double2 *)rc,
                     (double2 *)yc, active);
  HIP_LAUNCH_CHECK();
}

template <typename X>
__global__ void k_coarse_prolong(int64_t n, int nc, const int32_t *__restrict__ row_ptr,
                                 const int32_t *__restrict__ row_col, const double *__restrict__ row_val,
                                 const double2 *__restrict__ yc, X *__restrict__ z, const int32_t *__restrict__ active) {
  const int s = blockIdx.y, S = gridDim.y;
  if (active && !active[s]) return;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int b = row_ptr[i], e = row_ptr[i + 1];
    if (b == e) continue;
    double2 acc = to_d2(z[i * S + s]);
    for (int k = b; k < e; k++) {
      const double c = row_val[k];
      const double2 y = yc[(int64_t)s * nc + row_col[k]];
      acc.x += c * y.x;
      acc.y += c * y.y;
    }
    z[i * S + s] = from_d2<X>(acc);
  }
}

template <typename X>
void launch_coarse_prolong(hipStream_t st, int S, int64_t n, int nc, const int32_t *row_ptr, const int32_t *row_col,
                           const double *row_val, const double *yc, X *z, const int32_t *active) {
  int64_t bl = (n + 255) / 256;
  if (bl > 4096) bl = 4096;
  hipLaunchKernelGGL((k_coarse_prolong<X>), dim3((unsigned)bl, S), dim3(256), 0, st, n, nc, row_ptr, row_col, row_val,
                     (const double2 *)yc, z, active);
  HIP_LAUNCH_CHECK();
}
template void launch_coarse_prolong<double2>(hipStream_t, int, int64_t, int, const int32_t *, const int32_t *,
                                             const double *, const double *, double2 *, const int32_t *);
template void launch_coarse_prolong<float2>(hipStream_t, int, int64_t, int, const int32_t *, const int32_t *,
                                            const double *, const double *, float2 *, const int32_t *);

// ------------------------------------------------------------------ misc vector ops
template <typename X, typename Y>
__global__ void k_convert(int64_t count, const X *__restrict__ in, Y *__restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = from_d2<Y>(to_d2(in[i]));
}
template <typename X, typename Y>
void launch_convert(hipStream_t st, int64_t count, const X *in, Y *out) {
  int64_t bl = (count + 255) / 256;
  if (bl > 16384) bl = 16384;
  if (bl < 1) bl = 1;
  hipLaunchKernelGGL((k_convert<X, Y>), dim3((unsigned)bl), dim3(256), 0, st, count, in, out);
  HIP_LAUNCH_CHECK();
}
template void launch_convert<double2, float2>(hipStream_t, int64_t, const double2 *, float2 *);
template void launch_convert<float2, double2>(hipStream_t, int64_t, const float2 *, double2 *);
template void launch_convert<double2, double2>(hipStream_t, int64_t, const double2 *, double2 *);

template <typename X, typename Z>
__global__ void k_axpy(int64_t n, const Z *__restrict__ z, X *__restrict__ x, const int32_t *__restrict__ active) {
  const int s = blockIdx.y;
  if (active && !active[s]) return;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double2 a = to_d2(x[s * n + i]);
    const double2 b = to_d2(z[s * n + i]);
    x[s * n + i] = from_d2<X>(make_double2(a.x + b.x, a.y + b.y));
  }
}
template <typename X, typename Z>
void launch_axpy(hipStream_t st, int S, int64_t n, const Z *z, X *x, const int32_t *active) {
  int64_t bl = (n + 255) / 256;
  if (bl > 4096) bl = 4096;
  hipLaunchKernelGGL((k_axpy<X, Z>), dim3((unsigned)bl, S), dim3(256), 0, st, n, z, x, active);
  HIP_LAUNCH_CHECK();
}
template void launch_axpy<double2, double2>(hipStream_t, int, int64_t, const double2 *, double2 *, const int32_t *);
template void launch_axpy<double2, float2>(hipStream_t, int, int64_t, const float2 *, double2 *, const int32_t *);

template <typename X>
__global__ __launch_bounds__(kRedThreads) void k_norm2(int64_t n, const X *__restrict__ x, double *__restrict__ part,
                                                       int nblk, const int32_t *__restrict__ active) {
  const int s = blockIdx.y;
  if (active && !active[s]) return;
  __shared__ double red[kRedThreads / 64];
  const int64_t base = (int64_t)blockIdx.x * kChunk + threadIdx.x;
  double t = 0.0;
#pragma unroll
  for (int e = 0; e < kRedElems; e++) {
    const int64_t i = base + e * kRedThreads;
    if (i < n) {
      const double2 a = to_d2(x[s * n + i]);
      t += a.x * a.x + a.y * a.y;
    }
  }
  t = wave_sum(t);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = t;
  __syncthreads();
  if (threadIdx.x == 0) {
    double u = 0.0;
    for (int q = 0; q < kRedThreads / 64; q++) u += red[q];
    part[(int64_t)s * nblk + blockIdx.x] = u;
  }
}
template <typename X>
void launch_norm2(hipStream_t st, int S, int64_t n, const X *x, double *part, int nblk, const int32_t *active) {
  hipLaunchKernelGGL((k_norm2<X>), dim3(nblk, S), dim3(kRedThreads), 0, st, n, x, part, nblk, active);
  HIP_LAUNCH_CHECK();
}
template void launch_norm2<double2>(hipStream_t, int, int64_t, const double2 *, double *, int, const int32_t *);
template void launch_norm2<float2>(hipStream_t, int, int64_t, const float2 *, double *, int, const int32_t *);

// ------------------------------------------------------------------ Givens / Hessenberg
__global__ void k_givens(int S, int m, int j, int jlast, double2 *__restrict__ H, double *__restrict__ cs,
                         double2 *__restrict__ sn, double2 *__restrict__ g, double *__restrict__ est,
                         const double *__restrict__ thresh, int32_t *__restrict__ active,
                         int32_t *__restrict__ kfinal, const double *__restrict__ nrm2) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S || !active[s]) return;
  double2 *h = H + ((int64_t)s * m + j) * (m + 1);
  h[j + 1] = make_double2(sqrt(nrm2[s]), 0.0);
  double *c = cs + (int64_t)s * m;
  double2 *sv = sn + (int64_t)s * m;
  for (int i = 0; i < j; i++) {
    const double2 a = h[i], b = h[i + 1];
    const double2 sb = cmul(sv[i], b);
    const double2 t = make_double2(c[i] * a.x + sb.x, c[i] * a.y + sb.y);
    const double2 csa = cconjmul(sv[i], a);  // conj(s)*a
    h[i + 1] = make_double2(-csa.x + c[i] * b.x, -csa.y + c[i] * b.y);
    h[i] = t;
  }
  const double2 a = h[j], b = h[j + 1];
  const double na = sqrt(a.x * a.x + a.y * a.y), nb = sqrt(b.x * b.x + b.y * b.y);
  double cc;
  double2 ss, rr;
  if (nb == 0.0) {
    cc = 1.0;
    ss = make_double2(0.0, 0.0);
    rr = a;
  } else if (na == 0.0) {
    cc = 0.0;
    ss = make_double2(b.x / nb, -b.y / nb);  // conj(b)/|b|
    rr = make_double2(nb, 0.0);
  } else {
    const double nr = sqrt(na * na + nb * nb);
    cc = na / nr;
    const double2 ph = make_double2(a.x / na, a.y / na);
    ss = cmul(ph, make_double2(b.x / nr, -b.y / nr));  // (a/|a|) conj(b)/r
    rr = make_double2(ph.x * nr, ph.y * nr);
  }
  c[j] = cc;
  sv[j] = ss;
  h[j] = rr;
  h[j + 1] = make_double2(0.0, 0.0);
  double2 *gs = g + (int64_t)s * (m + 1);
  const double2 gj = gs[j];
  const double2 csg = cconjmul(ss, gj);
  gs[j + 1] = make_double2(-csg.x, -csg.y);
  gs[j] = make_double2(cc * gj.x, cc * gj.y);
  const double e = sqrt(gs[j + 1].x * gs[j + 1].x + gs[j + 1].y * gs[j + 1].y);
  est[s] = e;
  const bool breakdown = !(nrm2[s] > 0.0);
  if (e <= thresh[s] || breakdown || j >= jlast) {
    kfinal[s] = j + 1;
    active[s] = 0;
  }
}

void launch_givens_n(hipStream_t st, int S, int m, int j, int jlast, double *H, double *cs, double *sn, double *g,
                     double *est, const double *thresh, int32_t *active, int32_t *kfinal, const double *nrm2) {
  hipLaunchKernelGGL(k_givens, dim3((S + 63) / 64), dim3(64), 0, st, S, m, j, jlast, (double2 *)H, cs,
                     (double2 *)sn, (double2 *)g, est, thresh, active, kfinal, nrm2);
  HIP_LAUNCH_CHECK();
}

__global__ void k_backsolve(int S, int m, const double2 *__restrict__ H, const double2 *__restrict__ g,
                            const int32_t *__restrict__ kfinal, double2 *__restrict__ y,
                            const int32_t *__restrict__ cycle_active) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S || !cycle_active[s]) return;
  const int k = kfinal[s];
  const double2 *Hs = H + (int64_t)s * m * (m + 1);
  const double2 *gs = g + (int64_t)s * (m + 1);
  double2 *ys = y + (int64_t)s * m;
  for (int i = k - 1; i >= 0; i--) {
    double2 t = gs[i];
    for (int l = i + 1; l < k; l++) {
      const double2 u = cmul(Hs[(int64_t)l * (m + 1) + i], ys[l]);
      t.x -= u.x;
      t.y -= u.y;
    }
    const double2 d = Hs[(int64_t)i * (m + 1) + i];
    const double dm = d.x * d.x + d.y * d.y;
    ys[i] = dm > 0.0 ? cmul(t, make_double2(d.x / dm, -d.y / dm)) : make_double2(0.0, 0.0);
  }
  for (int i = k; i < m; i++) ys[i] = make_double2(0.0, 0.0);
}

void launch_backsolve(hipStream_t st, int S, int m, const double *H, const double *g, const int32_t *kfinal,
                      double *y, const int32_t *cycle_active) {
  hipLaunchKernelGGL(k_backsolve, dim3((S + 63) / 64), dim3(64), 0, st, S, m, (const double2 *)H,
                     (const double2 *)g, kfinal, (double2 *)y, cycle_active);
  HIP_LAUNCH_CHECK();
}

template <typename X>
__global__ void k_combine_basis(int64_t n, int m, const X *__restrict__ Vb, int64_t ldv_sys,
                                const double2 *__restrict__ y, const int32_t *__restrict__ kfinal, X *__restrict__ u,
                                const int32_t *__restrict__ cycle_active) {
  const int s = blockIdx.y;
  if (!cycle_active[s]) return;
  const int k = kfinal[s];
  __shared__ double2 ys[kMaxK];
  for (int i = threadIdx.x; i < k; i += blockDim.x) ys[i] = y[(int64_t)s * m + i];
  __syncthreads();
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double2 acc = make_double2(0.0, 0.0);
    for (int v = 0; v < k; v++) {
      const double2 c = cmul(to_d2(Vb[s * ldv_sys + (int64_t)v * n + i]), ys[v]);
      acc.x += c.x;
      acc.y += c.y;
    }
    u[s * n + i] = from_d2<X>(acc);
  }
}

template <typename X>
void launch_combine_basis(hipStream_t st, int S, int64_t n, int m, const X *Vb, int64_t ldv_sys, const double *y,
                          const int32_t *kfinal, X *u, const int32_t *cycle_active) {
  int64_t bl = (n + 255) / 256;
  if (bl > 4096) bl = 4096;
  hipLaunchKernelGGL((k_combine_basis<X>), dim3((unsigned)bl, S), dim3(256), 0, st, n, m, Vb, ldv_sys,
                     (const double2 *)y, kfinal, u, cycle_active);
  HIP_LAUNCH_CHECK();
}
template void launch_combine_basis<double2>(hipStream_t, int, int64_t, int, const double2 *, int64_t,
                                            const double *, const int32_t *, double2 *, const int32_t *);
template void launch_combine_basis<float2>(hipStream_t, int, int64_t, int, const float2 *, int64_t, const double *,
                                           const int32_t *, float2 *, const int32_t *);

// ------------------------------------------------------------------ port projections
__global__ void k_port_dots(int nport, const int32_t *__restrict__ pd_off, const int32_t *__restrict__ pd,
                            const double2 *__restrict__ u, int64_t ldu, const double2 *__restrict__ x, int64_t ldx,
                            double2 *__restrict__ out) {
  const int s = blockIdx.y, k = blockIdx.x;
  __shared__ double red[2][4];
  const int b0 = pd_off[k], b1 = pd_off[k + 1];
  double sx = 0.0, sy = 0.0;
  for (int i = b0 + threadIdx.x; i < b1; i += blockDim.x) {
    const double2 c = cconjmul(u[s * ldu + i], x[s * ldx + pd[i]]);
    sx += c.x;
    sy += c.y;
  }
  sx = wave_sum(sx);
  sy = wave_sum(sy);
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = sx;
    red[1][threadIdx.x >> 6] = sy;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double tx = 0.0, ty = 0.0;
    for (int q = 0; q < 4; q++) {
      tx += red[0][q];
      ty += red[1][q];
    }
    out[(int64_t)s * nport + k] = make_double2(tx, ty);
  }
}

void launch_port_dots(hipStream_t st, int S, int nport, const int32_t *pd_off, const int32_t *pd, const double *u,
                      int64_t ldu, const double *x, int64_t ldx, double *out) {
  hipLaunchKernelGGL(k_port_dots, dim3(nport, S), dim3(256), 0, st, nport, pd_off, pd, (const double2 *)u, ldu,
                     (const double2 *)x, ldx, (double2 *)out);
  HIP_LAUNCH_CHECK();
}

}  // namespace opem
