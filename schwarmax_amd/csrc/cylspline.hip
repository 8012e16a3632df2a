// CylSpline potential (CylindricalSpline.py:269-536, CubicSpline.py:102-305):
// the azimuthal-mode quadrature of the builder, the Legendre function
// Q_{m-1/2} it needs, and the spline evaluator of Phi, -grad Phi and the
// Poisson density at points.
#include <hip/hip_runtime.h>
#include <cmath>

#include "common.hpp"
#include "../../include/smx.h"

namespace smx {

constexpr double kPi = 3.14159265358979323846;
constexpr double kEuler = 0.57721566490153286061;
constexpr double kLn2 = 0.69314718055994530942;
// The reference evaluates Q's logarithmic term at max(1 - 1/chi^2, 1e-12)
// (CylindricalSpline.py:112).
constexpr double kQSwitch = 0.75;

// Q_{m-1/2}(chi) for chi = 1 + delta >= 1, delta given separately so that it
// keeps its relative accuracy near the singularity at chi = 1.  Toroidal form
//   Q_{m-1/2}(cosh eta) = sqrt(pi) Gamma(m+1/2) / Gamma(m+1) e^{-(m+1/2) eta}
//                         F(1/2, m+1/2; m+1; w),   w = e^{-2 eta},
// summed directly for w <= 0.75 (terms decrease monotonically), and for
// w > 0.75 by the logarithmic continuation of F(a, b; a+b; w) around w = 1
// (Abramowitz & Stegun 15.3.10), whose prefactor Gamma(a+b)/(Gamma(a)Gamma(b))
// cancels the one above:
//   Q = e^{-(m+1/2) eta} sum_n (a)_n (b)_n / (n!)^2
//       [2 psi(n+1) - psi(a+n) - psi(b+n) - ln(1-w)] (1-w)^n.
__device__ double legendre_q_half(int m, double delta) {
  const double chi = 1.0 + delta;
  const double y = delta * (2.0 + delta) / (chi * chi);     // 1 - 1/chi^2
  if (!(y >= 1e-12)) {
    // The reference's logarithm floor: in the continuation of
    // F((2m+1)/4, (2m+3)/4; m+1; 1 - y) around y = 0 (Abramowitz & Stegun
    // 15.3.10), ln y becomes ln 1e-12; terms past y^1 are below 1e-20.  The
    // prefactors sqrt(pi) Gamma(m+1/2)/(Gamma(m+1) 2^(m+1/2)) and
    // Gamma(m+1)/(Gamma(a) Gamma(b)) multiply to 1/2.
    const double a = 0.25 * (2 * m + 1), b = a + 0.5;
    double pa = -kEuler - 0.5 * kPi - 3.0 * kLn2;    // psi(1/4)
    double pb = -kEuler + 0.5 * kPi - 3.0 * kLn2;    // psi(3/4)
    if (m & 1) { const double s = pa; pa = pb; pb = s; }
    for (double q = (m & 1) ? 0.75 : 0.25; q < a - 0.1; q += 1.0) pa += 1.0 / q;
    for (double q = (m & 1) ? 0.25 : 0.75; q < b - 0.1; q += 1.0) pb += 1.0 / q;
    const double Lf = log(1e-12);
    const double t0 = -2.0 * kEuler - pa - pb - Lf;
    const double t1 = a * b * (2.0 * (1.0 - kEuler) - (pa + 1.0 / a) - (pb + 1.0 / b) - Lf) * y;
    return 0.5 * pow(chi, -(m + 0.5)) * (t0 + t1);
  }
  const double s2 = sqrt(delta * (2.0 + delta));
  const double eta = log1p(delta + s2);
  const double e = exp(-(m + 0.5) * eta);
  const double w = exp(-2.0 * eta);
  const double a = 0.5, b = m + 0.5, c = m + 1.0;
  if (w <= kQSwitch) {
    double pref = kPi;
    for (int k = 1; k <= m; ++k) pref *= (k - 0.5) / k;
    double t = 1.0, s = 1.0;
    for (int n = 0; n < 400; ++n) {
      t *= (a + n) * (b + n) / ((c + n) * (n + 1.0)) * w;
      s += t;
      if (fabs(t) <= 1e-17 * s) break;
    }
    return pref * e * s;
  }
  const double omw = -expm1(-2.0 * eta);
  const double L = log(omw);
  double psi_n1 = -kEuler;                 // psi(n + 1)
  double psi_a = -kEuler - 2.0 * kLn2;     // psi(1/2 + n)
  double psi_b = psi_a;                    // psi(m + 1/2 + n)
  for (int j = 0; j < m; ++j) psi_b += 1.0 / (0.5 + j);
  double cn = 1.0, yn = 1.0, s = 0.0;
  for (int n = 0; n < 400; ++n) {
    const double term = cn * (2.0 * psi_n1 - psi_a - psi_b - L) * yn;
    s += term;
    if (n > 0 && fabs(term) <= 1e-17 * fabs(s)) break;
    cn *= (a + n) * (b + n) / ((n + 1.0) * (n + 1.0));
    psi_n1 += 1.0 / (n + 1.0);
    psi_a += 1.0 / (a + n);
    psi_b += 1.0 / (b + n);
    yn *= omw;
  }
  return e * s;
}

// The mode-m Green's-function kernel Xi_m(R, z; R', z') of
// CylindricalSpline.py:162-190: Q_{m-1/2}(chi) / (pi sqrt(R R')),
// chi = (R^2 + R'^2 + (z - z')^2) / (2 R R'), and for R or R' below 1e-3
// the m = 0 point-mass kernel 1/|x - x'| (zero for m > 0).  chi - 1 is formed
// as ((R - R')^2 + (z - z')^2) / (2 R R'), free of cancellation.
__device__ __forceinline__ double xi_kernel(int m, double R, double z, double Rp, double zp) {
  const double dz = z - zp;
  if (R < 1e-3 || Rp < 1e-3) return m == 0 ? 1.0 / sqrt(R * R + Rp * Rp + dz * dz) : 0.0;
  const double dR = R - Rp;
  const double delta = (dR * dR + dz * dz) / (2.0 * R * Rp);
  return legendre_q_half(m, delta) / (kPi * sqrt(R * Rp));
}

// One workgroup per target (m, i, j): Phi_m(R_i, z_j) = -G sum_k w_k rho_m(k)
// [Xi_m(R_i, z_j; R'_k, z'_k) + Xi_m(R_i, z_j; R'_k, -z'_k)], where w_k
// already holds the Simpson weight, 2 pi R'_k and the map's Jacobian
// (CylindricalSpline.py:258-268, 318-330).  rho is (n_modes, 2, n_nodes),
// phi (n_modes, 2, nR, nZ).
constexpr int kPhiThreads = 256;

__global__ __launch_bounds__(kPhiThreads) void cylspline_phi_m_kernel(
    int n_modes, const double* __restrict__ R, int nR, const double* __restrict__ Z, int nZ,
    const double* __restrict__ node_R, const double* __restrict__ node_z,
    const double* __restrict__ node_w, int64_t n_nodes, const double* __restrict__ rho, double G,
    double* __restrict__ phi) {
  const int64_t t = blockIdx.x;
  const int m = static_cast<int>(t / ((int64_t)nR * nZ));
  const int i = static_cast<int>((t / nZ) % nR);
  const int j = static_cast<int>(t % nZ);
  const double Rt = R[i], zt = Z[j];
  const double* rre = rho + (int64_t)(2 * m) * n_nodes;
  const double* rim = rre + n_nodes;
  double sre = 0.0, sim = 0.0;
  for (int64_t k = threadIdx.x; k < n_nodes; k += kPhiThreads) {
    const double wk = node_w[k];
    const double a = rre[k], b = rim[k];
    if (wk == 0.0 || (a == 0.0 && b == 0.0)) continue;
    const double Rp = node_R[k], zp = node_z[k];
    const double xi = xi_kernel(m, Rt, zt, Rp, zp) + xi_kernel(m, Rt, zt, Rp, -zp);
    sre += wk * a * xi;
    sim += wk * b * xi;
  }
  for (int off = 32; off > 0; off >>= 1) {
    sre += __shfl_down(sre, off, 64);
    sim += __shfl_down(sim, off, 64);
  }
  __shared__ double red[2][kPhiThreads / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    red[0][wave] = sre;
    red[1][wave] = sim;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0, b = 0.0;
    for (int q = 0; q < kPhiThreads / 64; ++q) {
      a += red[0][q];
      b += red[1][q];
    }
    const int64_t plane = (int64_t)nR * nZ;
    phi[(int64_t)(2 * m) * plane + (int64_t)i * nZ + j] = -G * a;
    phi[(int64_t)(2 * m + 1) * plane + (int64_t)i * nZ + j] = -G * b;
  }
}

// Interval of x in the increasing grid g[0..n-1]: searchsorted(g, x) - 1
// clipped to [0, n-2] (CubicSpline.py:130-136).
__device__ __forceinline__ int spline_interval(const double* __restrict__ g, int n, double x) {
  int lo = 0, hi = n;  // first index with g[idx] >= x lies in [lo, hi]
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (g[mid] < x) lo = mid + 1; else hi = mid;
  }
  int i = lo - 1;
  return i < 0 ? 0 : (i > n - 2 ? n - 2 : i);
}

// Phi, -grad Phi and (optionally) nabla^2 Phi / (4 pi G) at points from the
// mode tables.  Per mode and part the separable interpolant of
// CubicSpline.py:202-305 gives the value and its u, v derivatives (and the
// second derivatives the Laplacian needs); the modes are summed as
// f_m (Re cos m phi - Im sin m phi), f_0 = 1, f_m = 2 (CylindricalSpline.py:
// 418-480); the chain rule follows CylindricalSpline.py:489-520.  Points
// outside the table give 0 for all three, as the reference's fill does.
__global__ __launch_bounds__(128) void cylspline_eval_kernel(smx_cylspline t,
                                                             const double* __restrict__ xyz,
                                                             int64_t n, double* __restrict__ phi,
                                                             double* __restrict__ acc,
                                                             double* __restrict__ rho) {
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= n) return;
  const double x = xyz[3 * p], y = xyz[3 * p + 1], z = xyz[3 * p + 2];
  const double R = sqrt(x * x + y * y);
  const double za = fabs(z);
  const double R0 = t.R0;
  const double U = asinh(R / R0), V = asinh(za / R0);
  const int nR = t.nR, nZ = t.nZ;
  const bool oob = (U < t.u[0]) || (U > t.u[nR - 1]) || (V < t.v[0]) || (V > t.v[nZ - 1]);
  if (oob) {
    if (phi) phi[p] = 0.0;
    if (acc) acc[3 * p] = acc[3 * p + 1] = acc[3 * p + 2] = 0.0;
    if (rho) rho[p] = 0.0;
    return;
  }
  const int i = spline_interval(t.u, nR, U);
  const int j = spline_interval(t.v, nZ, V);
  const double hx = t.u[i + 1] - t.u[i], hy = t.v[j + 1] - t.v[j];
  const double tx = (U - t.u[i]) / hx, ty = (V - t.v[j]) / hy;
  const double sx = 1.0 - tx, sy = 1.0 - ty;
  const double hx2 = hx * hx / 6.0, hy2 = hy * hy / 6.0;
  const double cx0 = sx * sx * sx - sx, cx1 = tx * tx * tx - tx;      // value weights of M
  const double dx0 = -3.0 * sx * sx + 1.0, dx1 = 3.0 * tx * tx - 1.0;  // d/dtx of those
  const double cy0 = sy * sy * sy - sy, cy1 = ty * ty * ty - ty;
  const double dy0 = -3.0 * sy * sy + 1.0, dy1 = 3.0 * ty * ty - 1.0;
  const int M = t.n_modes, K = t.n_modes_eval;
  const double* c00 = t.coef + ((int64_t)i * nZ + j) * M * 6;
  const double* c10 = t.coef + ((int64_t)(i + 1) * nZ + j) * M * 6;
  const double* c01 = t.coef + ((int64_t)i * nZ + j + 1) * M * 6;
  const double* c11 = t.coef + ((int64_t)(i + 1) * nZ + j + 1) * M * 6;
  // cos and sin of phi = atan2(y, x); on the axis atan2's signed zeros decide
  // (phi = pi for x = -0), as in the reference.
  double c1, s1;
  if (R > 0.0) {
    c1 = x / R;
    s1 = y / R;
  } else {
    const double a0 = atan2(y, x);
    c1 = cos(a0);
    s1 = sin(a0);
  }
  double cm = 1.0, sm = 0.0;
  double F = 0.0, Fu = 0.0, Fv = 0.0, Fp = 0.0, Fuu = 0.0, Fvv = 0.0, Fpp = 0.0;
  for (int m = 0; m < K; ++m) {
    double val[2], du[2], dv[2], duu[2], dvv[2];
#pragma unroll
    for (int part = 0; part < 2; ++part) {
      const double z00 = c00[6 * m + part], z10 = c10[6 * m + part];
      const double z01 = c01[6 * m + part], z11 = c11[6 * m + part];
      const double mx00 = c00[6 * m + 2 + part], mx10 = c10[6 * m + 2 + part];
      const double mx01 = c01[6 * m + 2 + part], mx11 = c11[6 * m + 2 + part];
      const double my00 = c00[6 * m + 4 + part], my10 = c10[6 * m + 4 + part];
      const double my01 = c01[6 * m + 4 + part], my11 = c11[6 * m + 4 + part];
      const double f0 = sx * z00 + tx * z10 + (cx0 * mx00 + cx1 * mx10) * hx2;
      const double f1 = sx * z01 + tx * z11 + (cx0 * mx01 + cx1 * mx11) * hx2;
      const double g0 = (z10 - z00) + (dx0 * mx00 + dx1 * mx10) * hx2;    // d f0 / d tx
      const double g1 = (z11 - z01) + (dx0 * mx01 + dx1 * mx11) * hx2;
      const double q0 = (sx * mx00 + tx * mx10) * hx * hx;                 // d2 f0 / d tx2
      const double q1 = (sx * mx01 + tx * mx11) * hx * hx;
      const double my0 = sx * my00 + tx * my10, my1 = sx * my01 + tx * my11;
      const double dmy0 = my10 - my00, dmy1 = my11 - my01;
      val[part] = sy * f0 + ty * f1 + (cy0 * my0 + cy1 * my1) * hy2;
      du[part] = (sy * g0 + ty * g1 + (cy0 * dmy0 + cy1 * dmy1) * hy2) / hx;
      dv[part] = ((f1 - f0) + (dy0 * my0 + dy1 * my1) * hy2) / hy;
      duu[part] = (sy * q0 + ty * q1) / (hx * hx);
      dvv[part] = (sy * my0 + ty * my1);
    }
    const double fm = m == 0 ? 1.0 : 2.0;
    const double re_c = val[0] * cm - val[1] * sm;
    F += fm * re_c;
    Fu += fm * (du[0] * cm - du[1] * sm);
    Fv += fm * (dv[0] * cm - dv[1] * sm);
    Fp += fm * (-m * val[0] * sm - m * val[1] * cm);
    Fuu += fm * (duu[0] * cm - duu[1] * sm);
    Fvv += fm * (dvv[0] * cm - dvv[1] * sm);
    Fpp += fm * (-(double)m * m * re_c);
    const double cn = cm * c1 - sm * s1;
    sm = sm * c1 + cm * s1;
    cm = cn;
  }
  if (phi) phi[p] = F;
  const double iR = 1.0 / sqrt(R * R + R0 * R0);           // du/dR
  const double iZ = 1.0 / sqrt(za * za + R0 * R0);         // dv/d|z|
  const double FR = Fu * iR;
  const double sz = z > 0.0 ? 1.0 : (z < 0.0 ? -1.0 : 0.0);
  const double Fz = Fv * iZ * sz;
  if (acc) {
    const double tiny = 1e-12;
    const double invR = R > tiny ? 1.0 / R : 0.0;
    const double dRdx = R > tiny ? x * invR : 0.0, dRdy = R > tiny ? y * invR : 0.0;
    const double dpdx = R > tiny ? -y * invR * invR : 0.0, dpdy = R > tiny ? x * invR * invR : 0.0;
    acc[3 * p] = -(FR * dRdx + Fp * dpdx);
    acc[3 * p + 1] = -(FR * dRdy + Fp * dpdy);
    acc[3 * p + 2] = -Fz;
  }
  if (rho) {
    // Laplacian in cylindrical coordinates with u = asinh(R/R0), v = asinh(|z|/R0);
    // d|z|/dz is sign(z), 0 at z = 0 as in the reference's autodiff of abs.
    const double FRR = Fuu * iR * iR - Fu * R * iR * iR * iR;
    const double Fzz = Fvv * iZ * iZ * sz * sz - Fv * za * iZ * iZ * iZ;
    const double lap = FRR + FR / R + Fpp / (R * R) + Fzz;
    rho[p] = lap / (4.0 * kPi * t.G);
  }
}

__global__ void legendre_q_kernel(int m, const double* __restrict__ delta, int64_t n,
                                  double* __restrict__ out) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k < n) out[k] = legendre_q_half(m, delta[k]);
}

}  // namespace smx

using namespace smx;

extern "C" {

int smx_cylspline_phi_m(int32_t n_modes, const double* R, int32_t nR, const double* Z, int32_t nZ,
                        const double* node_R, const double* node_z, const double* node_w,
                        int64_t n_nodes, const double* rho_nodes, double G, double* phi_m,
                        void* stream) {
  SMX_REQUIRE(n_modes >= 1 && n_modes <= SMX_CYLSPLINE_MAX_MODES && nR >= 1 && nZ >= 1 &&
                  n_nodes >= 0 && R && Z && phi_m && (n_nodes == 0 || (node_R && node_z && node_w &&
                  rho_nodes)),
              "smx_cylspline_phi_m: bad arguments");
  const int64_t blocks = (int64_t)n_modes * nR * nZ;
  hipLaunchKernelGGL(cylspline_phi_m_kernel, dim3(static_cast<unsigned>(blocks)), dim3(kPhiThreads),
                     0, as_stream(stream), (int)n_modes, R, (int)nR, Z, (int)nZ, node_R, node_z,
                     node_w, n_nodes, rho_nodes, G, phi_m);
  SMX_LAUNCHED();
  return 0;
}

int smx_cylspline_eval(const smx_cylspline* t, const double* xyz, int64_t n, double* phi,
                       double* acc, double* rho, void* stream) {
  SMX_REQUIRE(t && xyz && n >= 0 && t->u && t->v && t->coef && t->nR >= 2 && t->nZ >= 2 &&
                  t->n_modes >= 1 && t->n_modes_eval >= 1 && t->n_modes_eval <= t->n_modes &&
                  t->R0 > 0.0,
              "smx_cylspline_eval: bad arguments");
  if (n == 0) return 0;
  hipLaunchKernelGGL(cylspline_eval_kernel, dim3(blocks_for(n, 128)), dim3(128), 0,
                     as_stream(stream), *t, xyz, n, phi, acc, rho);
  SMX_LAUNCHED();
  return 0;
}

int smx_legendre_q_half(int32_t m, const double* delta, int64_t n, double* out, void* stream) {
  SMX_REQUIRE(m >= 0 && m < SMX_CYLSPLINE_MAX_MODES && delta && out && n >= 0,
              "smx_legendre_q_half: bad arguments");
  if (n == 0) return 0;
  hipLaunchKernelGGL(legendre_q_kernel, dim3(blocks_for(n, 128)), dim3(128), 0,
                     as_stream(stream), (int)m, delta, n, out);
  SMX_LAUNCHED();
  return 0;
}

}  // extern "C"
