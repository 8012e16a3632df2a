// Canonical SchwarMAX potential and stellar density as device functions.
//
// NFW halo (potentials.py:18-27), Miyamoto-Nagai disc (potentials.py:29-65),
// Dehnen & Aly T3 bar and V4 bulge as needle convolutions
// (dehnen_bar.py:104-157, 211-241, 269-278, 913-933, 1254-1260, 1318-1326,
// 1371-1383, 1425-1438, 1526-1543), summed as likelihood.py:13-71 sums them.
// Templated on the scalar type: T = double gives Phi or rho; T = Dual<3>
// seeded on (x, y, z) gives Phi and its exact gradient in one pass.
//
// The needle kernels A_n, C_n of all odd orders come out of one run of the
// reference's two-step recurrence per offset, instead of one run per order:
// the recurrence is the same sequence of operations, so every order's value is
// unchanged.
#pragma once
#include "dual.hpp"
#include "../../include/smx.h"

namespace smx {

constexpr double kEps = 1e-12;   // constants.py:11
constexpr double kPi = 3.141592653589793;

template <class T>
__device__ __forceinline__ void rotate3(const double* R, const double* o, T x, T y, T z,
                                        T& rx, T& ry, T& rz) {
  T dx = x - o[0], dy = y - o[1], dz = z - o[2];
  rx = (R[0] * dx + R[1] * dy) + R[2] * dz;
  ry = (R[3] * dx + R[4] * dy) + R[5] * dz;
  rz = (R[6] * dx + R[7] * dy) + R[8] * dz;
}

// potentials.py:18-27
template <class T>
__device__ __forceinline__ T nfw_phi(const smx_potential& p, T x, T y, T z) {
  T rx, ry, rz;
  rotate3(p.nfw_rot, p.nfw_origin, x, y, z, rx, ry, rz);
  T qx = rx / p.nfw_a, qy = ry / p.nfw_b, qz = rz / p.nfw_c;
  T r = dsqrt(((qx * qx + qy * qy) + qz * qz) + kEps);
  return (p.nfw_mGM * dlog(1.0 + r / p.nfw_Rs)) / (r + kEps);
}

// potentials.py:51-65
template <class T>
__device__ __forceinline__ T mn_phi(const smx_potential& p, T x, T y, T z) {
  T rx, ry, rz;
  rotate3(p.mn_rot, p.mn_origin, x, y, z, rx, ry, rz);
  rx = rx + kEps; ry = ry + kEps; rz = rz + kEps;
  T R = dsqrt(rx * rx + ry * ry);
  T u = p.mn_Rs + dsqrt(rz * rz + p.mn_Hs * p.mn_Hs);
  T d2 = R * R + u * u;
  return p.mn_mGM / dsqrt(d2);
}

// potentials.py:29-49
__device__ __forceinline__ double mn_rho(const smx_potential& p, double x, double y, double z) {
  double rx, ry, rz;
  rotate3(p.mn_rot, p.mn_origin, x, y, z, rx, ry, rz);
  rx += kEps; ry += kEps; rz += kEps;
  const double R = sqrt(rx * rx + ry * ry);
  const double beta = sqrt(rz * rz + p.mn_Hs * p.mn_Hs);
  const double sb = p.mn_Rs + beta;
  const double D2 = R * R + sb * sb;
  const double num = p.mn_Rs * R * R + (p.mn_Rs + 3.0 * beta) * (sb * sb);
  const double den = (beta * beta * beta) * pow(D2, 2.5);
  return p.mn_rho0 * (num / den);
}

// A_n and C_n = x A_n - B_n for n = 1, 3, ..., 2*NO-1 at abscissa xp
// (dehnen_bar.py:104-157 odd branch, 269-278).
template <class T, int NO>
__device__ __forceinline__ void acrn_odd(T xp, T u2, T* A, T* C) {
  T R2 = xp * xp + u2;
  T iR2 = 1.0 / R2;
  T r = dsqrt(R2);
  T Ax = 1.0 / r;
  T a = dlog(xp + r);
  T b = r;
  A[0] = a;
  C[0] = xp * a - b;
#pragma unroll
  for (int k = 1; k < NO; ++k) {
    const double cur = 2.0 * k - 1.0;
    T an = ((cur - 1.0) * a + xp * Ax) / (cur * u2);
    T bn = (-Ax) / cur;
    Ax = Ax * iR2;
    a = an;
    b = bn;
    A[k] = a;
    C[k] = xp * a - b;
  }
}

// Barred M/r^n for the odd orders n = 2*k0+1 .. 2*NO-1 (dehnen_bar.py:211-241,
// 928-933), or the axisymmetric M/r^n when L <= 0 (913-925, 1254-1260).
// Writes F[k] for k = k0 .. NO-1.
template <class T, int NO>
__device__ __forceinline__ void needle_odd(double M, double L, double gamma, T x, T y, T Z,
                                           int k0, T* F) {
  if (L > 0.0) {
    const double c1 = M * (1.0 - gamma) / (2.0 * L);
    const double c2 = M * gamma / (L * L);
    T u2 = y * y + Z * Z;
    T Ap[NO], Cp[NO], Am[NO], Cm[NO];
    acrn_odd<T, NO>(x + L, u2, Ap, Cp);
    acrn_odd<T, NO>(x - L, u2, Am, Cm);
    if (gamma != 0.0) {
      T A0[NO], C0[NO];
      acrn_odd<T, NO>(x, u2, A0, C0);
#pragma unroll
      for (int k = 0; k < NO; ++k)
        if (k >= k0) F[k] = c1 * (Ap[k] - Am[k]) + c2 * ((Cp[k] + Cm[k]) - 2.0 * C0[k]);
    } else {
#pragma unroll
      for (int k = 0; k < NO; ++k)
        if (k >= k0) F[k] = c1 * (Ap[k] - Am[k]);
    }
  } else {
    T iR2 = 1.0 / ((x * x + y * y) + Z * Z);
    T f = M * dsqrt(iR2);
#pragma unroll
    for (int k = 0; k < NO; ++k) {
      if (k >= k0) F[k] = f;
      f = f * iR2;
    }
  }
}

// dehnen_bar.py:1318-1326 (without the -G factor)
template <class T>
__device__ __forceinline__ T t3_ph(const smx_potential& p, T x, T y, T z) {
  const double a = p.t3_a, b = p.t3_b;
  T ze = dsqrt(z * z + b * b);
  T Z = ze + a;
  T aZ = a * Z;
  T F[3];
  needle_odd<T, 3>(p.t3_M, p.t3_L, p.t3_gamma, x, y, Z, 0, F);
  T ph = F[0];
  ph = ph + (a * (Z - a / 3.0)) * F[1];
  ph = ph + (aZ * aZ) * F[2];
  return ph;
}

// dehnen_bar.py:1371-1383 (without the -G factor)
template <class T>
__device__ __forceinline__ T v4_ph(const smx_potential& p, T x, T y, T z) {
  const double a = p.v4_a, b = p.v4_b;
  T ze = dsqrt(z * z + b * b);
  T Z = ze + a;
  T aZ = a * Z;
  const double hB2 = 0.5 * b * b;
  T F[5];
  needle_odd<T, 5>(p.v4_M, p.v4_L, p.v4_gamma, x, y, Z, 0, F);
  T ph = F[0];
  ph = ph + (a * (Z - 0.4 * a) + hB2) * F[1];
  ph = ph + ((3.0 * a) * ((0.2 * aZ) * ((Z + Z) - a) + hB2 * (Z - 0.2 * a))) * F[2];
  ph = ph + ((aZ * aZ) * (aZ + 6.0 * hB2)) * F[3];
  ph = ph + (((((7.0 * hB2) * ((aZ * aZ) * aZ)) * Z) / ze)) * F[4];
  return ph;
}

// likelihood.py:46-71: NFW + MN + T3 + V4
template <class T>
__device__ __forceinline__ T phi_total(const smx_potential& p, T x, T y, T z) {
  T h = nfw_phi(p, x, y, z);
  T m = mn_phi(p, x, y, z);
  T t = (-p.G) * t3_ph(p, x, y, z);
  T v = (-p.G) * v4_ph(p, x, y, z);
  return ((h + m) + t) + v;
}

// Phi and a = -grad Phi (model.py:152-157)
__device__ __forceinline__ double accel(const smx_potential& p, double x, double y, double z,
                                        double& ax, double& ay, double& az) {
  Dual<3> X = dconst<3>(x), Y = dconst<3>(y), Zc = dconst<3>(z);
  X.g[0] = 1.0;
  Y.g[1] = 1.0;
  Zc.g[2] = 1.0;
  Dual<3> ph = phi_total(p, X, Y, Zc);
  ax = -ph.g[0];
  ay = -ph.g[1];
  az = -ph.g[2];
  return ph.v;
}

// dehnen_bar.py:1425-1438
__device__ __forceinline__ double t3_rho(const smx_potential& p, double x, double y, double z) {
  const double a = p.t3_a, b = p.t3_b;
  const double ze = sqrt(z * z + b * b);
  const double iz = 1.0 / ze;
  const double Z = ze + a;
  const double Z2 = Z * Z, A2 = a * a;
  double F[5];
  needle_odd<double, 5>(p.t3_M, p.t3_L, p.t3_gamma, x, y, Z, 2, F);
  double rh = (3.0 * ((ze * ze) * ze)) * F[2];
  rh = rh + (((5.0 * a) * Z2) * (((3.0 * ze) - (2.0 * a)) * ze + A2)) * F[3];
  rh = rh + ((((35.0 * A2) * Z2) * Z2) * ze) * F[4];
  return (((b * b) * ((iz * iz) * iz)) * (0.25 / kPi)) * rh;
}

// dehnen_bar.py:1526-1543
__device__ __forceinline__ double v4_rho(const smx_potential& p, double x, double y, double z) {
  const double a = p.v4_a, b = p.v4_b;
  const double ze = sqrt(z * z + b * b);
  const double iz = 1.0 / ze;
  const double Z = ze + a;
  const double Z2 = Z * Z, A2 = a * a, AZ = a * Z;
  double F[7];
  needle_odd<double, 7>(p.v4_M, p.v4_L, p.v4_gamma, x, y, Z, 3, F);
  double rh = (5.0 * ((ze * ze) * ze)) * F[3];
  rh = rh + (((7.0 * a) * Z2) * (((((5.0 * ze) * ze - (4.0 * a) * ze) + 3.0 * A2) -
                                  ((2.0 * a) * A2) * iz) + ((A2 * A2) * iz) * iz)) * F[4];
  rh = rh + ((((63.0 * A2) * Z2) * Z2) * (((ze + ze) - 2.0 * a) + A2 * iz)) * F[5];
  rh = rh + ((((((231.0 * A2) * AZ) * Z2) * Z2) * Z)) * F[6];
  return (((b * b) * ((iz * iz) * iz)) * (((0.375 * b) * b) / kPi)) * rh;
}

// likelihood.py:13-43: MN + T3 + V4
__device__ __forceinline__ double rho_total(const smx_potential& p, double x, double y, double z) {
  return (mn_rho(p, x, y, z) + t3_rho(p, x, y, z)) + v4_rho(p, x, y, z);
}

}  // namespace smx
