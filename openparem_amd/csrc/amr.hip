// Element errors of the ZZ flux-recovery estimate (fem3D::calculateMeshErrors, src/OpenParEM3D/fem3D.cpp:1162-1380,
// with OPEM_L2ZZErrorEstimator.cpp:18-164): for each system and element K,
//   e_re = int_K |Re(s - flux)| dV,  e_im = int_K |Im(s - flux)| dV,  e = sqrt(e_re^2 + e_im^2),
// flux = curl(H)/mu_r on K and s the ND_p L2 projection of the flux (solved by the caller), and the largest e over
// the systems (the driven ports) is kept per element.
//
// GPU layout: one wavefront per element, its 64 lanes over the quadrature points.  The element's gradients
// grad(lambda_m), the six cross products grad(lambda_i) x grad(lambda_j) and one system's H and s DOFs sit in LDS;
// each lane sums the basis at its points from per-(point, function) tables of lambda^alpha and its partial
// derivatives, then the wave adds the weighted norms with cross-lane shuffles.  The basis is the one of
// nd_tables.h: phi = lambda^alpha w, w = lambda_i grad(lambda_j) - lambda_j grad(lambda_i),
// curl(phi) = grad(lambda^alpha) x w + 2 lambda^alpha grad(lambda_i) x grad(lambda_j).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <vector>

#include "devbuf.h"
#include "nd_tables.h"
#include "quadrature.h"

namespace opem {

namespace {

constexpr int kZzWaves = 4;     // elements per workgroup
constexpr int kZzMaxNd = 45;    // ND3

struct ZzTabs {
  int nq, nd;
  const double *L;     // nq*4
  const double *w;     // nq
  const double *la;    // nq*nd
  const double *dla;   // nq*nd*4
  const int *fn;       // nd*4: i, j, pair index, unused
};

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__global__ __launch_bounds__(64 * kZzWaves) void k_zz_error(int64_t ne, int S, int64_t n, const double *__restrict__ coords,
                                                           const int32_t *__restrict__ edofs,
                                                           const double *__restrict__ inv_mur, const double2 *__restrict__ H,
                                                           const double2 *__restrict__ Sf, ZzTabs T,
                                                           double *__restrict__ err, double *__restrict__ err_sys) {
  __shared__ double sg[kZzWaves][4][3];
  __shared__ double sx[kZzWaves][6][3];
  __shared__ double2 sh[kZzWaves][kZzMaxNd], ss[kZzWaves][kZzMaxNd];
  const int wv = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int64_t e = (int64_t)blockIdx.x * kZzWaves + wv;
  const bool live = e < ne;
  const int nd = T.nd;
  // gradients of the barycentrics: rows of J^-1 for J = [v1-v0, v2-v0, v3-v0]
  double adet = 0.0;
  if (live && lane == 0) {
    const double *v = coords + e * 12;
    double J[3][3];
    for (int r = 0; r < 3; r++)
      for (int c = 0; c < 3; c++) J[r][c] = v[3 * (c + 1) + r] - v[r];
    const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                       J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                       J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    const double id = 1.0 / det;
    // inverse rows: grad lambda_{1..3}
    double Gi[3][3];
    Gi[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) * id;
    Gi[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
    Gi[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
    Gi[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) * id;
    Gi[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
    Gi[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
    Gi[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) * id;
    Gi[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
    Gi[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
    for (int d = 0; d < 3; d++) {
      sg[wv][1][d] = Gi[0][d];
      sg[wv][2][d] = Gi[1][d];
      sg[wv][3][d] = Gi[2][d];
      sg[wv][0][d] = -(Gi[0][d] + Gi[1][d] + Gi[2][d]);
    }
    const int pr[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
    for (int q = 0; q < 6; q++) {
      const double *a = sg[wv][pr[q][0]], *b = sg[wv][pr[q][1]];
      sx[wv][q][0] = a[1] * b[2] - a[2] * b[1];
      sx[wv][q][1] = a[2] * b[0] - a[0] * b[2];
      sx[wv][q][2] = a[0] * b[1] - a[1] * b[0];
    }
  }
  if (live) {
    const double *v = coords + e * 12;
    double J00 = v[3] - v[0], J01 = v[6] - v[0], J02 = v[9] - v[0];
    double J10 = v[4] - v[1], J11 = v[7] - v[1], J12 = v[10] - v[1];
    double J20 = v[5] - v[2], J21 = v[8] - v[2], J22 = v[11] - v[2];
    adet = std::fabs(J00 * (J11 * J22 - J12 * J21) - J01 * (J10 * J22 - J12 * J20) + J02 * (J10 * J21 - J11 * J20));
  }
  const double imu = live ? inv_mur[e] : 0.0;
  double emax = 0.0;
  for (int sys = 0; sys < S; sys++) {
    __syncthreads();
    if (live && lane < nd) {
      const int64_t g = edofs[e * nd + lane];
      sh[wv][lane] = H[(int64_t)sys * n + g];
      ss[wv][lane] = Sf[(int64_t)sys * n + g];
    }
    __syncthreads();
    double ere = 0.0, eim = 0.0;
    if (live) {
      for (int q = lane; q < T.nq; q += 64) {
        const double L0 = T.L[4 * q], L1 = T.L[4 * q + 1], L2 = T.L[4 * q + 2], L3 = T.L[4 * q + 3];
        const double Lq[4] = {L0, L1, L2, L3};
        double s_re[3] = {0, 0, 0}, s_im[3] = {0, 0, 0}, c_re[3] = {0, 0, 0}, c_im[3] = {0, 0, 0};
        for (int a = 0; a < nd; a++) {
          const int i = T.fn[4 * a], j = T.fn[4 * a + 1], pq = T.fn[4 * a + 2];
          const double la = T.la[(int64_t)q * nd + a];
          const double *dl = T.dla + ((int64_t)q * nd + a) * 4;
          const double Li = Lq[i], Lj = Lq[j];
          double wv3[3], gl[3], cu[3];
#pragma unroll
          for (int d = 0; d < 3; d++) {
            wv3[d] = Li * sg[wv][j][d] - Lj * sg[wv][i][d];
            gl[d] = dl[0] * sg[wv][0][d] + dl[1] * sg[wv][1][d] + dl[2] * sg[wv][2][d] + dl[3] * sg[wv][3][d];
          }
          cu[0] = gl[1] * wv3[2] - gl[2] * wv3[1] + 2.0 * la * sx[wv][pq][0];
          cu[1] = gl[2] * wv3[0] - gl[0] * wv3[2] + 2.0 * la * sx[wv][pq][1];
          cu[2] = gl[0] * wv3[1] - gl[1] * wv3[0] + 2.0 * la * sx[wv][pq][2];
          const double2 hv = sh[wv][a], sv = ss[wv][a];
#pragma unroll
          for (int d = 0; d < 3; d++) {
            const double ph = la * wv3[d];
            s_re[d] += ph * sv.x;
            s_im[d] += ph * sv.y;
            c_re[d] += cu[d] * hv.x;
            c_im[d] += cu[d] * hv.y;
          }
        }
        double nr = 0.0, ni = 0.0;
#pragma unroll
        for (int d = 0; d < 3; d++) {
          const double dr = s_re[d] - imu * c_re[d], di = s_im[d] - imu * c_im[d];
          nr += dr * dr;
          ni += di * di;
        }
        const double wq = T.w[q] * adet;
        ere += wq * std::sqrt(nr);
        eim += wq * std::sqrt(ni);
      }
    }
    ere = wave_sum(ere);
    eim = wave_sum(eim);
    const double es = std::sqrt(ere * ere + eim * eim);
    emax = sys == 0 ? es : std::fmax(emax, es);
    if (live && lane == 0 && err_sys) err_sys[e * S + sys] = es;
  }
  if (live && lane == 0) err[e] = emax;
}

}  // namespace

// Element errors on the stream st: coords (ne*12, sorted vertices), edofs (device, ne*nd), inv_mur (ne), H and Sf
// (device, S systems x n complex, system-major).  err (host, ne), err_sys (host, ne*S, optional).
void zz_errors_run(hipStream_t st, int p, int64_t ne, int S, int64_t n, const std::vector<double> &coords,
                   const int32_t *d_edofs, const std::vector<double> &inv_mur, const double2 *d_H, const double2 *d_S,
                   double *err, double *err_sys) {
  const std::vector<BasisFn> basis = nd_tet_basis(p);
  const int nd = (int)basis.size();
  if (nd > kZzMaxNd) throw std::invalid_argument("order above 3");
  std::vector<double> L, w;
  tet_rule(p + 2, L, w);
  const int nq = (int)w.size();
  std::vector<double> la((size_t)nq * nd), dla((size_t)nq * nd * 4);
  std::vector<int> fn((size_t)nd * 4, 0);
  const int pair_id[4][4] = {{-1, 0, 1, 2}, {0, -1, 3, 4}, {1, 3, -1, 5}, {2, 4, 5, -1}};
  for (int a = 0; a < nd; a++) {
    fn[4 * a] = basis[a].i;
    fn[4 * a + 1] = basis[a].j;
    fn[4 * a + 2] = pair_id[basis[a].i][basis[a].j];
  }
  for (int q = 0; q < nq; q++)
    for (int a = 0; a < nd; a++) {
      const double *Lq = &L[4 * q];
      double v = 1.0;
      for (int m = 0; m < 4; m++) v *= std::pow(Lq[m], basis[a].alpha[m]);
      la[(size_t)q * nd + a] = v;
      for (int m = 0; m < 4; m++) {
        const int am = basis[a].alpha[m];
        double d = 0.0;
        if (am > 0) {
          d = am;
          for (int k = 0; k < 4; k++) d *= std::pow(Lq[k], basis[a].alpha[k] - (k == m ? 1 : 0));
        }
        dla[((size_t)q * nd + a) * 4 + m] = d;
      }
    }
  DevBuf dL, dw, dla_, dd, dfn, dc, dmu, derr, dsys;
  dL.upload(L, st);
  dw.upload(w, st);
  dla_.upload(la, st);
  dd.upload(dla, st);
  dfn.upload(fn, st);
  dc.upload(coords, st);
  dmu.upload(inv_mur, st);
  derr.alloc((size_t)ne * 8);
  if (err_sys) dsys.alloc((size_t)ne * S * 8);
  const ZzTabs T{nq, nd, dL.as<double>(), dw.as<double>(), dla_.as<double>(), dd.as<double>(), dfn.as<int>()};
  const int64_t blocks = (ne + kZzWaves - 1) / kZzWaves;
  hipLaunchKernelGGL(k_zz_error, dim3((unsigned)blocks), dim3(64 * kZzWaves), 0, st, ne, S, n, dc.as<double>(),
                     d_edofs, dmu.as<double>(), d_H, d_S, T, derr.as<double>(),
                     err_sys ? dsys.as<double>() : nullptr);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(err, derr.p, (size_t)ne * 8, hipMemcpyDeviceToHost, st));
  if (err_sys) HIPCHK(hipMemcpyAsync(err_sys, dsys.p, (size_t)ne * S * 8, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
}

}  // namespace opem
