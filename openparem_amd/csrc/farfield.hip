// Far field of radiation-boundary currents and the pattern sphere/circle geometry and metrics.
//
// Far field (BoundaryDatabase::calculateFarField, src/OpenParEM3D/port.cpp:8077-8160; Balanis, Advanced
// Engineering Electromagnetics, 1989, pp. 288-289): for far points P (relative to a centre) and surface currents
// J, M with areas dA at r',
//   N = sum J dA e^{j beta rhat.r'},  L = sum M dA e^{j beta rhat.r'},
//   E_theta = -k (L_phi + eta N_theta),  E_phi = k (L_theta - eta N_phi),
//   H_theta =  k (N_phi - L_theta / eta), H_phi = -k (N_theta + L_phi / eta),  k = j beta e^{-j beta r} / (4 pi r).
// The reference projects each current onto theta/phi inside the loop; the projection is linear, so here the
// Cartesian N and L are summed first and projected once per point.
//
// GPU layout: currents in HBM as 16 doubles each (r' relative to the centre, J dA and M dA interleaved complex,
// one pad), points as 4 doubles (rhat, r).  A workgroup of 256 threads takes 256 points and one chunk of the
// currents, staged through LDS 64 currents at a time; each thread sums its point's 12 partial sums in registers
// and writes them once.  Chunks are sized so that the launch has at least ~2048 workgroups whatever the number of
// points; a second kernel adds the chunks in a fixed order (deterministic) and applies the far-field factors.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "devbuf.h"

namespace opem {

namespace {

constexpr int kFfThreads = 256;
constexpr int kFfTile = 64;       // currents per LDS stage
constexpr int kCurStride = 16;    // doubles per current
constexpr int kPartStride = 12;   // doubles per (chunk, point) partial: N (3 complex), L (3 complex)

__global__ __launch_bounds__(kFfThreads) void k_far_field_part(int64_t npts, const double4 *__restrict__ pts,
                                                               int64_t ncur, int64_t chunk,
                                                               const double *__restrict__ cur, double beta,
                                                               double *__restrict__ part) {
  __shared__ double sC[kFfTile * kCurStride];
  const int64_t p = (int64_t)blockIdx.x * kFfThreads + threadIdx.x;
  const int64_t c0 = (int64_t)blockIdx.y * chunk;
  const int64_t c1 = std::min<int64_t>(ncur, c0 + chunk);
  double4 P = p < npts ? pts[p] : make_double4(0.0, 0.0, 1.0, 1.0);
  const double bx = beta * P.x, by = beta * P.y, bz = beta * P.z;
  double acc[12];
#pragma unroll
  for (int q = 0; q < 12; q++) acc[q] = 0.0;
  for (int64_t t0 = c0; t0 < c1; t0 += kFfTile) {
    const int nt = (int)std::min<int64_t>(kFfTile, c1 - t0);
    __syncthreads();
    for (int q = threadIdx.x; q < nt * kCurStride; q += kFfThreads) sC[q] = cur[t0 * kCurStride + q];
    __syncthreads();
    for (int k = 0; k < nt; k++) {
      const double *c = sC + k * kCurStride;
      const double ph = bx * c[0] + by * c[1] + bz * c[2];
      double s, co;
      sincos(ph, &s, &co);
#pragma unroll
      for (int q = 0; q < 6; q++) {
        const double re = c[3 + 2 * q], im = c[4 + 2 * q];
        acc[2 * q] += co * re - s * im;
        acc[2 * q + 1] += co * im + s * re;
      }
    }
  }
  if (p < npts) {
    double *o = part + ((int64_t)blockIdx.y * npts + p) * kPartStride;
#pragma unroll
    for (int q = 0; q < 12; q++) o[q] = acc[q];
  }
}

__global__ void k_far_field_sum(int64_t npts, int nchunk, const double4 *__restrict__ pts,
                                const double *__restrict__ part, double beta, double eta,
                                double *__restrict__ out) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npts) return;
  double a[12];
#pragma unroll
  for (int q = 0; q < 12; q++) a[q] = 0.0;
  for (int ch = 0; ch < nchunk; ch++) {
    const double *o = part + ((int64_t)ch * npts + p) * kPartStride;
#pragma unroll
    for (int q = 0; q < 12; q++) a[q] += o[q];
  }
  const double4 P = pts[p];
  // direction cosines from rhat: cos theta = z, sin theta = rho, cos phi = x / rho, sin phi = y / rho
  const double rho = std::sqrt(P.x * P.x + P.y * P.y);
  const double ct = P.z, st = rho;
  const double cp = rho > 0 ? P.x / rho : 1.0, sp = rho > 0 ? P.y / rho : 0.0;
  // components: N = a[0..5], L = a[6..11] (x re, x im, y re, y im, z re, z im)
  double Nt[2], Np[2], Lt[2], Lp[2];
  for (int r = 0; r < 2; r++) {
    Nt[r] = a[r] * ct * cp + a[2 + r] * ct * sp - a[4 + r] * st;
    Np[r] = -a[r] * sp + a[2 + r] * cp;
    Lt[r] = a[6 + r] * ct * cp + a[8 + r] * ct * sp - a[10 + r] * st;
    Lp[r] = -a[6 + r] * sp + a[8 + r] * cp;
  }
  // k = j beta e^{-j beta r} / (4 pi r) = beta / (4 pi r) (sin(beta r) + j cos(beta r))
  double sr, cr;
  sincos(beta * P.w, &sr, &cr);
  const double m = beta / (4.0 * M_PI * P.w);
  const double kr = m * sr, ki = m * cr;
  auto kmul = [&](double xr, double xi, double *o) {
    o[0] = kr * xr - ki * xi;
    o[1] = kr * xi + ki * xr;
  };
  double *o = out + p * 8;
  // E_phi, E_theta, H_phi, H_theta
  kmul(Lt[0] - eta * Np[0], Lt[1] - eta * Np[1], o + 0);
  kmul(-(Lp[0] + eta * Nt[0]), -(Lp[1] + eta * Nt[1]), o + 2);
  kmul(-(Nt[0] + Lp[0] / eta), -(Nt[1] + Lp[1] / eta), o + 4);
  kmul(Np[0] - Lt[0] / eta, Np[1] - Lt[1] / eta, o + 6);
}

}  // namespace

// Host side of the far-field evaluation on the stream st (synchronous: returns with out filled).
// xyz: currents' positions (ncur*3), J, M: complex ncur*3, area: ncur; pts: npts*3; out: complex npts*4.
// kernel_ms (optional): the two kernels' time from HIP events.
void far_field_run(hipStream_t st, int64_t ncur, const double *xyz, const double *J, const double *M,
                   const double *area, const double *center, double beta, double eta, int64_t npts,
                   const double *pts, double *out, double *kernel_ms) {
  if (kernel_ms) *kernel_ms = 0.0;
  if (npts == 0) return;
  if (ncur == 0) {  // no currents: no far field
    std::memset(out, 0, (size_t)npts * 64);
    return;
  }
  std::vector<double> hc((size_t)ncur * kCurStride, 0.0);
  for (int64_t k = 0; k < ncur; k++) {
    double *c = hc.data() + k * kCurStride;
    for (int d = 0; d < 3; d++) c[d] = xyz[3 * k + d] - center[d];
    for (int d = 0; d < 3; d++) {
      c[3 + 2 * d] = J[6 * k + 2 * d] * area[k];
      c[4 + 2 * d] = J[6 * k + 2 * d + 1] * area[k];
      c[9 + 2 * d] = M[6 * k + 2 * d] * area[k];
      c[10 + 2 * d] = M[6 * k + 2 * d + 1] * area[k];
    }
  }
  std::vector<double> hp((size_t)npts * 4);
  for (int64_t p = 0; p < npts; p++) {
    const double x = pts[3 * p] - center[0], y = pts[3 * p + 1] - center[1], z = pts[3 * p + 2] - center[2];
    const double r = std::sqrt(x * x + y * y + z * z);
    if (!(r > 0)) throw std::invalid_argument("a far-field point coincides with the centre");
    hp[4 * p] = x / r;
    hp[4 * p + 1] = y / r;
    hp[4 * p + 2] = z / r;
    hp[4 * p + 3] = r;
  }
  const int64_t pblocks = (npts + kFfThreads - 1) / kFfThreads;
  int64_t nchunk = std::max<int64_t>(1, std::min<int64_t>((ncur + kFfTile - 1) / kFfTile,
                                                           (2048 + pblocks - 1) / pblocks));
  nchunk = std::min<int64_t>(nchunk, 65535);
  int64_t chunk = (ncur + nchunk - 1) / nchunk;
  chunk = ((chunk + kFfTile - 1) / kFfTile) * kFfTile;
  nchunk = std::max<int64_t>(1, (ncur + chunk - 1) / chunk);
  DevBuf dc, dp, dpart, dout;
  dc.upload(hc, st);
  dp.upload(hp, st);
  dpart.alloc((size_t)nchunk * npts * kPartStride * 8);
  dout.alloc((size_t)npts * 64);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (kernel_ms) {
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e0, st));
  }
  hipLaunchKernelGGL(k_far_field_part, dim3((unsigned)pblocks, (unsigned)nchunk), dim3(kFfThreads), 0, st, npts,
                     dp.as<double4>(), ncur, chunk, dc.as<double>(), beta, dpart.as<double>());
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(k_far_field_sum, dim3((unsigned)((npts + 255) / 256)), dim3(256), 0, st, npts, (int)nchunk,
                     dp.as<double4>(), dpart.as<double>(), beta, eta, dout.as<double>());
  HIPCHK(hipGetLastError());
  if (kernel_ms) {
    HIPCHK(hipEventRecord(e1, st));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    *kernel_ms = ms;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
  }
  HIPCHK(hipMemcpyAsync(out, dout.p, (size_t)npts * 64, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
}

namespace {

// spherical excess of the triangle of unit directions a, b, c (Van Oosterom and Strackee, IEEE Trans. Biomed.
// Eng. 30, 1983): tan(E/2) = |a.(b x c)| / (1 + a.b + b.c + c.a)
double excess(const double *a, const double *b, const double *c) {
  const double bxc[3] = {b[1] * c[2] - b[2] * c[1], b[2] * c[0] - b[0] * c[2], b[0] * c[1] - b[1] * c[0]};
  const double num = std::fabs(a[0] * bxc[0] + a[1] * bxc[1] + a[2] * bxc[2]);
  const double den = 1.0 + (a[0] * b[0] + a[1] * b[1] + a[2] * b[2]) + (b[0] * c[0] + b[1] * c[1] + b[2] * c[2]) +
                     (c[0] * a[0] + c[1] * a[1] + c[2] * a[2]);
  return 2.0 * std::atan2(num, den);
}

void unit_from(const double *p, const double *c, double *u) {
  const double d[3] = {p[0] - c[0], p[1] - c[1], p[2] - c[2]};
  const double n = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  for (int k = 0; k < 3; k++) u[k] = d[k] / n;
}

struct SphereGeo {
  std::vector<double> pts;       // vertices then triangle centres, 3 each
  std::vector<int32_t> tris;     // 3 each
  std::vector<double> tri_area;
  std::vector<double> point_area;
  int64_t nvert = 0;
};

// Sphere::create, refine, setMetrics and allocateAreasToPoints (src/OpenParEM3D/pattern.cpp:618-800): the
// icosahedron about the origin, each refinement replacing triangle (a, b, c) in place by (a, ab, ca) and appending
// (ab, b, bc), (ab, bc, ca), (ca, bc, c), with edge midpoints pushed to the sphere and shared through an edge map;
// then one centre per triangle (its vertices' mean about the centre, pushed to the sphere), triangle areas and
// point areas (each vertex takes the two sub-triangles vertex-midpoint-centre at its corner) about the centre.
void build_sphere(double radius, const double *center, int refine, SphereGeo &g) {
  static const double X = 0.525731112119133606, Z = 0.850650808352039932;
  const double v0[12][3] = {{-X, 0, Z}, {X, 0, Z}, {-X, 0, -Z}, {X, 0, -Z}, {0, Z, X}, {0, Z, -X},
                            {0, -Z, X}, {0, -Z, -X}, {Z, X, 0}, {-Z, X, 0}, {Z, -X, 0}, {-Z, -X, 0}};
  const int t0[20][3] = {{0, 4, 1}, {0, 9, 4}, {9, 5, 4}, {4, 5, 8}, {4, 8, 1}, {8, 10, 1}, {8, 3, 10},
                         {5, 3, 8}, {5, 2, 3}, {2, 7, 3}, {7, 10, 3}, {7, 6, 10}, {7, 11, 6}, {11, 0, 6},
                         {0, 1, 6}, {6, 1, 10}, {9, 0, 11}, {9, 11, 2}, {9, 2, 5}, {7, 2, 11}};
  std::vector<double> &P = g.pts;
  P.clear();
  for (auto &v : v0)
    for (int d = 0; d < 3; d++) P.push_back(v[d] * radius);
  std::vector<int32_t> &T = g.tris;
  T.assign(&t0[0][0], &t0[0][0] + 60);
  for (int it = 0; it < refine; it++) {
    std::unordered_map<int64_t, int32_t> mids;
    auto mid = [&](int32_t i, int32_t j) {
      const int64_t key = (int64_t)std::min(i, j) << 32 | (int64_t)std::max(i, j);
      auto f = mids.find(key);
      if (f != mids.end()) return f->second;
      double m[3];
      for (int d = 0; d < 3; d++) m[d] = 0.5 * (P[3 * i + d] + P[3 * j + d]);
      const double len = std::sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
      for (int d = 0; d < 3; d++) P.push_back(m[d] / (len / radius));
      const int32_t id = (int32_t)(P.size() / 3 - 1);
      mids.emplace(key, id);
      return id;
    };
    const size_t n0 = T.size() / 3;
    for (size_t t = 0; t < n0; t++) {
      const int32_t a = T[3 * t], b = T[3 * t + 1], c = T[3 * t + 2];
      const int32_t ab = mid(a, b), bc = mid(b, c), ca = mid(c, a);
      const int32_t add[9] = {ab, b, bc, ab, bc, ca, ca, bc, c};
      T.insert(T.end(), add, add + 9);
      T[3 * t + 1] = ab;
      T[3 * t + 2] = ca;
    }
  }
  g.nvert = (int64_t)P.size() / 3;
  const size_t nt = T.size() / 3;
  g.tri_area.assign(nt, 0.0);
  const double r2 = radius * radius;
  for (size_t t = 0; t < nt; t++) {
    const double *A = &P[3 * T[3 * t]], *B = &P[3 * T[3 * t + 1]], *C = &P[3 * T[3 * t + 2]];
    double m[3];
    for (int d = 0; d < 3; d++) m[d] = ((A[d] - center[d]) + (B[d] - center[d]) + (C[d] - center[d])) / 3.0;
    const double len = std::sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
    double ua[3], ub[3], uc[3];
    unit_from(A, center, ua);
    unit_from(B, center, ub);
    unit_from(C, center, uc);
    g.tri_area[t] = r2 * excess(ua, ub, uc);
    for (int d = 0; d < 3; d++) P.push_back(m[d] / (len / radius) + center[d]);
  }
  g.point_area.assign(P.size() / 3, 0.0);
  for (size_t t = 0; t < nt; t++) {
    const int32_t ia = T[3 * t], ib = T[3 * t + 1], ic = T[3 * t + 2];
    const double *A = &P[3 * ia], *B = &P[3 * ib], *C = &P[3 * ic];
    double mab[3], mbc[3], mca[3], u[3][3], ctr[3], w[3];
    unit_from(&P[3 * (g.nvert + (int64_t)t)], center, ctr);
    for (int d = 0; d < 3; d++) {
      w[d] = 0.5 * (A[d] + B[d]);
    }
    unit_from(w, center, mab);
    for (int d = 0; d < 3; d++) w[d] = 0.5 * (B[d] + C[d]);
    unit_from(w, center, mbc);
    for (int d = 0; d < 3; d++) w[d] = 0.5 * (C[d] + A[d]);
    unit_from(w, center, mca);
    unit_from(A, center, u[0]);
    unit_from(B, center, u[1]);
    unit_from(C, center, u[2]);
    g.point_area[ia] += r2 * (excess(u[0], mab, ctr) + excess(u[0], ctr, mca));
    g.point_area[ib] += r2 * (excess(u[1], mbc, ctr) + excess(u[1], ctr, mab));
    g.point_area[ic] += r2 * (excess(u[2], mca, ctr) + excess(u[2], ctr, mbc));
  }
}

}  // namespace

}  // namespace opem

using namespace opem;

extern "C" {

int opem_pattern_sphere(double radius, const double *center, int refine, int64_t *npts, double *pts,
                        double *point_area, int64_t *ntri, int32_t *tris, double *tri_area) {
  if (!(radius > 0) || !center || refine < 0 || refine > 9 || !npts || !ntri) return -2;
  try {
    SphereGeo g;
    build_sphere(radius, center, refine, g);
    *npts = (int64_t)g.pts.size() / 3;
    *ntri = (int64_t)g.tris.size() / 3;
    if (pts) std::memcpy(pts, g.pts.data(), g.pts.size() * 8);
    if (point_area) std::memcpy(point_area, g.point_area.data(), g.point_area.size() * 8);
    if (tris) std::memcpy(tris, g.tris.data(), g.tris.size() * 4);
    if (tri_area) std::memcpy(tri_area, g.tri_area.data(), g.tri_area.size() * 8);
    return 0;
  } catch (...) {
    return 1;
  }
}

int opem_pattern_circle(double radius, const double *center, double theta, double phi, double latitude,
                        int64_t npoints, double *pts) {
  if (!(radius > 0) || !center || npoints < 1 || !pts) return -2;
  const double del = 2 * M_PI / (double)npoints;
  double ang = 0.0;
  for (int64_t i = 0; i < npoints; i++) {
    const double x = radius * std::cos(ang) * std::cos(latitude) + center[0];
    const double y = radius * std::sin(ang) * std::cos(latitude) + center[1];
    const double z = radius * std::sin(latitude) + center[2];
    // rotate_point (pattern.cpp:1302-1310): theta about y, then phi about z, about the origin
    const double xr = std::cos(theta) * x + std::sin(theta) * z;
    const double zr = -std::sin(theta) * x + std::cos(theta) * z;
    pts[3 * i] = std::cos(phi) * xr - std::sin(phi) * y;
    pts[3 * i + 1] = std::sin(phi) * xr + std::cos(phi) * y;
    pts[3 * i + 2] = zr;
    ang += del;
  }
  return 0;
}

int opem_pattern_metrics(int64_t npts, const double *fields, const double *point_area, double total_area,
                         const double *accepted_power, double *radiated_power, double *gain, double *directivity,
                         double *summary) {
  if (npts < 1 || !fields || !point_area || !(total_area > 0) || !accepted_power || !radiated_power || !summary)
    return -2;
  typedef struct {
    double re, im;
  } c2;
  auto cmul = [](c2 a, c2 b) { return c2{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; };
  auto cdiv = [](c2 a, c2 b) {
    const double d = b.re * b.re + b.im * b.im;
    return c2{(a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d};
  };
  std::vector<c2> pw((size_t)npts);
  c2 prad{0.0, 0.0};
  for (int64_t i = 0; i < npts; i++) {
    const double *f = fields + 8 * i;  // E_phi, E_theta, H_phi, H_theta
    const c2 a = cmul(c2{f[2], f[3]}, c2{f[4], -f[5]});
    const c2 b = cmul(c2{f[0], f[1]}, c2{f[6], -f[7]});
    pw[i] = c2{0.5 * (a.re - b.re), 0.5 * (a.im - b.im)};
    prad.re += pw[i].re * point_area[i];
    prad.im += pw[i].im * point_area[i];
  }
  radiated_power[0] = prad.re;
  radiated_power[1] = prad.im;
  const c2 acc{accepted_power[0], accepted_power[1]};
  const c2 accd{acc.re / total_area, acc.im / total_area}, radd{prad.re / total_area, prad.im / total_area};
  double gmax = -HUGE_VAL, dmax = -HUGE_VAL;
  for (int64_t i = 0; i < npts; i++) {
    const double gi = cdiv(pw[i], accd).re, di = cdiv(pw[i], radd).re;
    if (gain) gain[i] = gi;
    if (directivity) directivity[i] = di;
    gmax = std::max(gmax, 10 * std::log10(gi));
    dmax = std::max(dmax, 10 * std::log10(di));
  }
  summary[0] = gmax;
  summary[1] = dmax;
  summary[2] = cdiv(prad, acc).re;
  return 0;
}

int opem_accepted_power(int nports, const double *Cp, const double *Cm, const double *Pz, int sport,
                        double *accepted) {
  if (nports < 1 || !Cp || !Cm || !Pz || sport < 0 || sport >= nports || !accepted) return -2;
  auto at = [&](const double *A, int i, int k, int r) { return A[2 * ((int64_t)i * nports + k) + r]; };
  const double pr = Pz[0], pi = Pz[1];
  const double cin = at(Cm, sport, sport, 0) * at(Cm, sport, sport, 0) + at(Cm, sport, sport, 1) * at(Cm, sport, sport, 1);
  double re = cin * pr, im = cin * pi;
  for (int i = 0; i < nports; i++) {
    if (i == sport) continue;
    const double co = at(Cp, i, sport, 0) * at(Cp, i, sport, 0) + at(Cp, i, sport, 1) * at(Cp, i, sport, 1);
    re -= co * pr;
    im -= co * pi;
  }
  accepted[0] = re;
  accepted[1] = im;
  return 0;
}

}  // extern "C"
