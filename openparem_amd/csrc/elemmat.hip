// Element matrices of independent affine tets on MFMA (gfx950).
//
// For one element, A_e = K_e - k0^2 er (1 - j tand) M_e with
//   K_e = sum_c gk_c TK_c,  M_e = sum_c gm_c TM_c      (6 symmetric Jacobian components c)
// so a tile of 16 elements is a GEMM  X[16 x K] * T[K x nd^2]:
//   Re: X = [gk(6), -k0^2 er gm(6), 0(4)],   T = [TK; TM; 0]     (K = 16)
//   Im: X = [k0^2 er tand gm(6), 0(2)],       T = [TM; 0]         (K = 8)
// One wavefront owns 16 elements and walks the nd^2 output in 16-wide tiles.
//   prec 0: v_mfma_f64_16x16x4f64, FP64 out
//   prec 1: v_mfma_f32_16x16x4f32, FP32 out (factors formed in FP64, rounded once)
//   prec 2: v_mfma_f32_16x16x32_bf16 with per-element scaling of X to max|X| = 1, FP32 out
// Operand layouts (CDNA3/4 16x16 MFMA): A[i][k] in lane i + 16*(k/kpl), B[k][j] in lane j + 16*(k/kpl),
// D[i][j] in register i%4 of lane j + 16*(i/4) for f32 outputs, register i/4 of lane j + 16*(i%4) for
// v_mfma_f64_16x16x4f64; kpl = values per lane (1 for f64/f32, 8 for bf16).
#include <hip/hip_runtime.h>

#include <cstring>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "kernels.h"
#include "nd_tables.h"

namespace opem {

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));
typedef __bf16 bf8_t __attribute__((ext_vector_type(8)));

static inline int padded(int nd) { return ((nd * nd + 15) / 16) * 16; }

// geometric factors of one tet in FP64 (same formulas as tet_factors, device side)
__device__ void dev_factors(const double *v, double gm[6], double gk[6]) {
  double J[3][3];
  for (int c = 0; c < 3; c++)
    for (int r = 0; r < 3; r++) J[r][c] = v[(c + 1) * 3 + r] - v[r];
  const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                     J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                     J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
  const double id = 1.0 / det;
  double inv[3][3];
  inv[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) * id;
  inv[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
  inv[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
  inv[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) * id;
  inv[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
  inv[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
  inv[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) * id;
  inv[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
  inv[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
  const int comp[6][2] = {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {0, 2}, {1, 2}};
  const double ad = fabs(det);
  for (int c = 0; c < 6; c++) {
    const int k = comp[c][0], l = comp[c][1];
    double G = 0.0, C = 0.0;
    for (int m = 0; m < 3; m++) {
      G += inv[k][m] * inv[l][m];
      C += J[m][k] * J[m][l];
    }
    gm[c] = ad * G;
    gk[c] = C / ad;
  }
}

template <int PREC>
struct ElemT {
  using T = double;  // table / A-operand type
  using O = double;  // output type
};
template <>
struct ElemT<1> {
  using T = float;
  using O = float;
};
template <>
struct ElemT<2> {
  using T = __bf16;
  using O = float;
};

// tab layout: Tre[16][NP] then Tim[8][NP] (K rows, padded N); bf16: Tre[32][NP], Tim[32][NP]
// Output staging: a wave's 16 elements own one contiguous output range (16 x 2 nd^2 values), but an MFMA tile
// holds only 16 columns of each element, i.e. 64-128 byte pieces that straddle cache lines whenever nd^2 is not a
// multiple of 16.  The wave therefore computes CW columns (CW/16 tiles) of both planes into its LDS stage
// Stage[wave][2][16][CW] and then stores each element's CW-long runs with consecutive lanes on consecutive addresses.
template <int PREC>
struct ElemCW {
  static constexpr int CW = PREC == 0 ? 32 : 64;  // 8 KB of stage per wave at every precision
};

// NT: non-temporal output stores.  They pay when every run covers whole 64-byte segments: element blocks (nd^2
// values) that are whole numbers of segments (ND2: +5-40% measured) and the contiguous WHOLE range; where runs
// straddle segments (ND3) the L2 must merge the partial lines and plain stores are faster.
// WHOLE (ND1 only, nd^2 = 36): a wave's 16 element blocks (16 x 2 x 36 values, one contiguous output range) fit in
// its stage, so the tiles are staged in output order and the range is streamed out with consecutive lanes, with
// non-temporal stores (ND1 FP64 0.17 -> 0.12 ms per 1M elements measured against plain stores).
constexpr int kWholeNdd = 36;
// PLANE (ND3): one plane (Re, then Im) of 16 elements at a time, in runs twice as long as the two-plane chunks, for
// the long misaligned element blocks of ND3 (nd^2 = 2025).
template <int PREC, bool NT, int MODE>  // MODE 0: two-plane chunks, 1: WHOLE, 2: PLANE
__global__ __launch_bounds__(256) void k_elem(int64_t ne, int nd, int NP, const double *__restrict__ verts,
                                              const double *__restrict__ er, const double *__restrict__ tand,
                                              double k0, const typename ElemT<PREC>::T *__restrict__ tab,
                                              typename ElemT<PREC>::O *__restrict__ out) {
  using T = typename ElemT<PREC>::T;
  using O = typename ElemT<PREC>::O;
  constexpr int CW = ElemCW<PREC>::CW;
  // Per wave one 8 KB region: first the 16 elements' factors X[16][24] (16 re + 8 im) and BF16 scales sc[16],
  // read into registers before the column loop, then the output stage [2][16][CW].
  constexpr bool WHOLE = MODE == 1, PLANE = MODE == 2;
  constexpr int CWS = PLANE ? 2 * CW : CW;                          // staged columns per chunk
  constexpr int SW = WHOLE ? 16 * 2 * kWholeNdd : 2 * 16 * CW;      // stage values per wave
  constexpr int TPC = WHOLE ? (kWholeNdd + 15) / 16 : CWS / 16;     // column tiles per staged chunk
  static_assert(SW * sizeof(O) >= 16 * 24 * 8 + 16 * 8, "stage too small for the factors");
  __shared__ __align__(16) O Stage[4][SW];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double(*X1)[24] = (double(*)[24])(void *)Stage[wave];
  double *sc1 = (double *)(void *)Stage[wave] + 16 * 24;
  const int64_t e0 = ((int64_t)blockIdx.x * 4 + wave) * 16;
  if (lane < 16) {
    const int64_t e = e0 + lane;
    double row[24];
    for (int k = 0; k < 24; k++) row[k] = 0.0;
    if (e < ne) {
      double gm[6], gk[6];
      dev_factors(verts + e * 12, gm, gk);
      const double ke = k0 * k0 * er[e];
      for (int c = 0; c < 6; c++) {
        row[c] = gk[c];
        row[6 + c] = -ke * gm[c];
        row[16 + c] = ke * tand[e] * gm[c];
      }
    }
    double s = 1.0;
    if (PREC == 2) {
      double mr = 0.0, mi = 0.0;
      for (int k = 0; k < 16; k++) mr = fmax(mr, fabs(row[k]));
      for (int k = 16; k < 24; k++) mi = fmax(mi, fabs(row[k]));
      s = mr > 0.0 ? mr : 1.0;
      const double si = mi > 0.0 ? mi : 1.0;
      for (int k = 0; k < 16; k++) row[k] /= s;
      for (int k = 16; k < 24; k++) row[k] /= si;
      sc1[lane] = s;
      row[23] = si;  // imag scale in the pad slot (row[22..23] are zero padding)
    }
    for (int k = 0; k < 24; k++) X1[lane][k] = row[k];
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  if (e0 >= ne) return;
  const int i16 = lane & 15, kq = lane >> 4;
  const int ndd = nd * nd;
  const int nel = ne - e0 < 16 ? (int)(ne - e0) : 16;
  O *ob = out + e0 * (int64_t)(2 * ndd);
  const int ntile = NP / 16;
  O(*st)[16][CW] = (O(*)[16][CW])(void *)Stage[wave];
  O *wst = Stage[wave];
  // A operands, fixed over the column tiles
  T are[4], aim[2];
  bf8_t bre_a, bim_a;
  if constexpr (PREC == 0 || PREC == 1) {
    for (int kb = 0; kb < 4; kb++) are[kb] = (T)X1[i16][kb * 4 + kq];
    for (int kb = 0; kb < 2; kb++) aim[kb] = (T)X1[i16][16 + kb * 4 + kq];
  } else {
    // bf16: one K=32 step each; lanes kq = 0,1 carry K rows 0..15, kq = 2,3 zeros
    for (int t = 0; t < 8; t++) {
      const int k = kq * 8 + t;
      bre_a[t] = (__bf16)(k < 16 ? (float)X1[i16][k] : 0.0f);
      bim_a[t] = (__bf16)(k < 6 ? (float)X1[i16][16 + k] : 0.0f);
    }
  }
  double scr[4] = {1.0, 1.0, 1.0, 1.0}, sci[4] = {1.0, 1.0, 1.0, 1.0};
  if constexpr (PREC == 2) {
    for (int r = 0; r < 4; r++) {
      scr[r] = sc1[kq * 4 + r];
      sci[r] = X1[kq * 4 + r][23];
    }
  }
  // the factors are in registers: the region becomes the output stage
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  for (int part = 0; part < (PLANE ? 2 : 1); part++)
  for (int c0 = 0; c0 < ndd; c0 += WHOLE ? ndd : CWS) {
#pragma unroll
    for (int tt = 0; tt < TPC; tt++) {
      const int t = c0 / 16 + tt;
      if (t >= ntile) break;
      const int n = t * 16 + i16;
      f4_t fre = {0, 0, 0, 0}, fim = {0, 0, 0, 0};
      d4_t dre = {0, 0, 0, 0}, dim = {0, 0, 0, 0};
      if constexpr (PREC == 0 || PREC == 1) {
        const T *Tre = tab;
        const T *Tim = tab + 16 * NP;
        if (!PLANE || part == 0) {
#pragma unroll
          for (int kb = 0; kb < 4; kb++) {
            const T b = Tre[(kb * 4 + kq) * NP + n];
            if constexpr (PREC == 0)
              dre = __builtin_amdgcn_mfma_f64_16x16x4f64(are[kb], b, dre, 0, 0, 0);
            else
              fre = __builtin_amdgcn_mfma_f32_16x16x4f32(are[kb], b, fre, 0, 0, 0);
          }
        }
        if (!PLANE || part == 1) {
#pragma unroll
          for (int kb = 0; kb < 2; kb++) {
            const T b = Tim[(kb * 4 + kq) * NP + n];
            if constexpr (PREC == 0)
              dim = __builtin_amdgcn_mfma_f64_16x16x4f64(aim[kb], b, dim, 0, 0, 0);
            else
              fim = __builtin_amdgcn_mfma_f32_16x16x4f32(aim[kb], b, fim, 0, 0, 0);
          }
        }
      } else {
        const __bf16 *Tre = tab;
        const __bf16 *Tim = tab + 32 * NP;
        bf8_t bre, bim;
        for (int q = 0; q < 8; q++) {
          bre[q] = Tre[(kq * 8 + q) * NP + n];
          bim[q] = Tim[(kq * 8 + q) * NP + n];
        }
        fre = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bre_a, bre, fre, 0, 0, 0);
        fim = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bim_a, bim, fim, 0, 0, 0);
      }
#pragma unroll
      for (int r = 0; r < 4; r++) {
        // f64 16x16x4: D row i in lane group i%4, register i/4; f32 16x16x4: group i/4, register i%4
        const int el = PREC == 0 ? r * 4 + kq : kq * 4 + r;
        O vre, vim;
        if constexpr (PREC == 0) {
          vre = dre[r];
          vim = dim[r];
        } else {
          vre = (O)(fre[r] * scr[r]);
          vim = (O)(fim[r] * sci[r]);
        }
        if constexpr (WHOLE) {
          if (n < ndd) {  // output order: element, plane, column
            wst[el * 2 * ndd + n] = vre;
            wst[el * 2 * ndd + ndd + n] = vim;
          }
        } else if constexpr (PLANE) {
          wst[el * CWS + tt * 16 + i16] = part == 0 ? vre : vim;
        } else {
          st[0][el][tt * 16 + i16] = vre;
          st[1][el][tt * 16 + i16] = vim;
        }
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if constexpr (WHOLE) {
      const int total = nel * 2 * ndd;  // the wave's contiguous output range
      for (int i = lane; i < total; i += 64) {
        if constexpr (NT)
          __builtin_nontemporal_store(wst[i], ob + i);
        else
          ob[i] = wst[i];
      }
      continue;  // one pass covers every column
    }
    if constexpr (PLANE) {
      const int len = ndd - c0 < CWS ? ndd - c0 : CWS;
#pragma unroll
      for (int it = 0; it < 16 * CWS / 64; it++) {
        const int idx = it * 64 + lane;
        const int col = idx % CWS, el = idx / CWS;
        if (col < len && el < nel) {
          O *dst = ob + (int64_t)el * 2 * ndd + part * ndd + c0 + col;
          if constexpr (NT)
            __builtin_nontemporal_store(wst[el * CWS + col], dst);
          else
            *dst = wst[el * CWS + col];
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      continue;
    }
    const int len = ndd - c0 < CW ? ndd - c0 : CW;
#pragma unroll
    for (int it = 0; it < 2 * 16 * CW / 64; it++) {
      const int idx = it * 64 + lane;
      const int col = idx % CW, rest = idx / CW;
      const int el = rest & 15, part = rest >> 4;
      if (col < len && el < nel) {
        O *dst = ob + (int64_t)el * 2 * ndd + part * ndd + c0 + col;
        if constexpr (NT)
          __builtin_nontemporal_store(st[part][el][col], dst);
        else
          *dst = st[part][el][col];
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

size_t element_tab_bytes(int prec, int p) {
  const int nd = nd_tet_ndof(p);
  const int NP = padded(nd);
  if (prec == 0) return (size_t)24 * NP * 8;
  if (prec == 1) return (size_t)24 * NP * 4;
  return (size_t)64 * NP * 2;
}

size_t element_geo_bytes(int, int64_t) { return 0; }

// Host: build the padded tables in the layout the kernel expects.
static std::vector<unsigned char> make_tab(int prec, int p) {
  const TetTables &T = tet_tables(p);
  const int nd = T.nd, ndd = nd * nd, NP = padded(nd);
  const int kre = prec == 2 ? 32 : 16, kim = prec == 2 ? 32 : 8;
  std::vector<double> re((size_t)kre * NP, 0.0), im((size_t)kim * NP, 0.0);
  for (int c = 0; c < 6; c++)
    for (int n = 0; n < ndd; n++) {
      re[(size_t)c * NP + n] = T.K[(size_t)c * ndd + n];
      re[(size_t)(6 + c) * NP + n] = T.M[(size_t)c * ndd + n];
      im[(size_t)c * NP + n] = T.M[(size_t)c * ndd + n];
    }
  std::vector<unsigned char> out(element_tab_bytes(prec, p));
  size_t k = 0;
  auto put = [&](const std::vector<double> &v) {
    for (double x : v) {
      if (prec == 0) {
        std::memcpy(&out[k], &x, 8);
        k += 8;
      } else if (prec == 1) {
        float f = (float)x;
        std::memcpy(&out[k], &f, 4);
        k += 4;
      } else {
        __bf16 b = (__bf16)(float)x;
        std::memcpy(&out[k], &b, 2);
        k += 2;
      }
    }
  };
  put(re);
  put(im);
  return out;
}

void element_table(int prec, int p, void *host_out) {
  auto t = make_tab(prec, p);
  std::memcpy(host_out, t.data(), t.size());
}

void launch_element_matrices(hipStream_t st, int prec, int p, int64_t ne, const double *verts, const double *er,
                             const double *tand, double k0, const void *tab, void *, void *out) {
  const int nd = nd_tet_ndof(p), NP = padded(nd);
  const unsigned blocks = (unsigned)((ne + 63) / 64);
  const bool whole = nd * nd <= kWholeNdd;
  // ND3: one plane at a time (measured per 1M elements against two-plane chunks: FP64 13.37 -> 11.63 ms, FP32
  // 7.52 -> 6.95 ms with non-temporal stores; BF16 6.85 -> 6.44 ms with plain stores, 6.69 ms non-temporal)
  const bool plane = nd * nd > 1000;
  const bool plane_nt = prec != 2;
  const bool nt = !whole && (nd * nd * (prec == 0 ? 8 : 4)) % 64 == 0;
#define OPEM_ELEM(P, TT, OO)                                                                                       \
  {                                                                                                              \
    if (whole)                                                                                                   \
      hipLaunchKernelGGL((k_elem<P, true, 1>), dim3(blocks), dim3(256), 0, st, ne, nd, NP, verts, er, tand, k0,    \
                         (const TT *)tab, (OO *)out);                                                            \
    else if (plane && plane_nt)                                                                                  \
      hipLaunchKernelGGL((k_elem<P, true, 2>), dim3(blocks), dim3(256), 0, st, ne, nd, NP, verts, er, tand, k0,    \
                         (const TT *)tab, (OO *)out);                                                            \
    else if (plane)                                                                                              \
      hipLaunchKernelGGL((k_elem<P, false, 2>), dim3(blocks), dim3(256), 0, st, ne, nd, NP, verts, er, tand, k0,   \
                         (const TT *)tab, (OO *)out);                                                            \
    else if (nt)                                                                                                 \
      hipLaunchKernelGGL((k_elem<P, true, 0>), dim3(blocks), dim3(256), 0, st, ne, nd, NP, verts, er, tand, k0,    \
                         (const TT *)tab, (OO *)out);                                                            \
    else                                                                                                         \
      hipLaunchKernelGGL((k_elem<P, false, 0>), dim3(blocks), dim3(256), 0, st, ne, nd, NP, verts, er, tand, k0,   \
                         (const TT *)tab, (OO *)out);                                                            \
  }
  if (prec == 0) {
    OPEM_ELEM(0, double, double)
  } else if (prec == 1) {
    OPEM_ELEM(1, float, float)
  } else if (prec == 2) {
    OPEM_ELEM(2, __bf16, float)
  } else
    throw std::runtime_error("element matrices: bad precision");
#undef OPEM_ELEM
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) throw std::runtime_error(std::string("element kernel launch: ") + hipGetErrorString(e));
}

}  // namespace opem
