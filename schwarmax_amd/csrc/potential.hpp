// Canonical SchwarMAX potential and stellar density as device functions.
//
// NFW halo (potentials.py:18-27), Miyamoto-Nagai disc (potentials.py:29-65),
// Dehnen & Aly T3 bar and V4 bulge as needle convolutions
// (dehnen_bar.py:104-157, 211-241, 269-278, 913-933, 1254-1260, 1318-1326,
// 1371-1383, 1425-1438, 1526-1543), summed as likelihood.py:13-71 sums them.
// The needle kernels A_n, C_n of all odd orders come out of one run of the
// reference's two-step recurrence per offset, instead of one run per order:
// the recurrence is the same sequence of operations, so every order's value is
// unchanged.  Values are folded into the convolution sums as each order is
// produced, which keeps few registers live.
#pragma once
#include <hip/hip_runtime.h>
#include "../../include/smx.h"

namespace smx {

constexpr double kEps = 1e-12;   // constants.py:11
constexpr double kPi = 3.141592653589793;

__device__ __forceinline__ void rotate3(const double* R, const double* o, double x, double y,
                                        double z, double& rx, double& ry, double& rz) {
  const double dx = x - o[0], dy = y - o[1], dz = z - o[2];
  rx = (R[0] * dx + R[1] * dy) + R[2] * dz;
  ry = (R[3] * dx + R[4] * dy) + R[5] * dz;
  rz = (R[6] * dx + R[7] * dy) + R[8] * dz;
}

// potentials.py:18-27
__device__ __forceinline__ double nfw_phi(const smx_potential& p, double x, double y, double z) {
  double rx, ry, rz;
  rotate3(p.nfw_rot, p.nfw_origin, x, y, z, rx, ry, rz);
  const double qx = rx / p.nfw_a, qy = ry / p.nfw_b, qz = rz / p.nfw_c;
  const double r = sqrt(((qx * qx + qy * qy) + qz * qz) + kEps);
  return (p.nfw_mGM * log(1.0 + r / p.nfw_Rs)) / (r + kEps);
}

// potentials.py:51-65
__device__ __forceinline__ double mn_phi(const smx_potential& p, double x, double y, double z) {
  double rx, ry, rz;
  rotate3(p.mn_rot, p.mn_origin, x, y, z, rx, ry, rz);
  rx = rx + kEps; ry = ry + kEps; rz = rz + kEps;
  const double R = sqrt(rx * rx + ry * ry);
  const double u = p.mn_Rs + sqrt(rz * rz + p.mn_Hs * p.mn_Hs);
  const double d2 = R * R + u * u;
  return p.mn_mGM / sqrt(d2);
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

// Run the reference's recurrence (dehnen_bar.py:104-157, odd branch; C_n =
// x A_n - B_n, 269-278) at abscissa xp and fold each order into running sums:
//   F[k] += sA * A_k + sC * C_k        (and, with the gradient, Q[k] += sA * Ax_k + sC * A_k)
// sA, sC in {0, +-1, -2}: the sums reproduce the reference's differences
// (A(x+L) - A(x-L)) and ((C(x+L) + C(x-L)) - 2 C(x)) operation for operation.
// EXACT divides by cur*u2 as the reference does; otherwise multiplies by a
// precomputed 1/u2 (force path).
template <int NO, bool EXACT, bool GRAD>
__device__ __forceinline__ void acrn_fold(double xp, double u2, double iu2, double sA, double sC,
                                          double* F, double* Q) {
  const double R2 = xp * xp + u2;
  const double iR2 = 1.0 / R2;
  const double r = sqrt(R2);
  double ax = 1.0 / r;
  double a = log(xp + r);
  double b = r;
#pragma unroll
  for (int k = 0; k < NO; ++k) {
    if (k > 0) {
      const double cur = 2.0 * k - 1.0;
      const double num = (cur - 1.0) * a + xp * ax;
      const double an = EXACT ? num / (cur * u2) : num * (iu2 / cur);
      b = (-ax) / cur;
      ax = ax * iR2;
      a = an;
    }
    if (sC == 0.0) {
      F[k] += sA * a;
      if (GRAD) Q[k] += sA * ax;
    } else {
      F[k] += sC * (xp * a - b);
      if (GRAD) Q[k] += sC * a;
    }
  }
}

// Needle-convolved M/r^n for n = 1, 3, .., 2*NO-1 (dehnen_bar.py:211-241,
// 928-933), or the axisymmetric M/r^n when L <= 0 (913-925, 1254-1260); with
// GRAD also dF/dx.
template <int NO, bool EXACT, bool GRAD>
__device__ __forceinline__ void needle(double M, double L, double gamma, double x, double u2,
                                       double* F, double* Fx) {
#pragma unroll
  for (int k = 0; k < NO; ++k) {
    F[k] = 0.0;
    if (GRAD) Fx[k] = 0.0;
  }
  if (L > 0.0) {
    const double c1 = M * (1.0 - gamma) / (2.0 * L);
    const double c2 = M * gamma / (L * L);
    const double iu2 = EXACT ? 0.0 : 1.0 / u2;
    if (gamma == 0.0) {
      acrn_fold<NO, EXACT, GRAD>(x + L, u2, iu2, 1.0, 0.0, F, Fx);
      acrn_fold<NO, EXACT, GRAD>(x - L, u2, iu2, -1.0, 0.0, F, Fx);
#pragma unroll
      for (int k = 0; k < NO; ++k) {
        F[k] = c1 * F[k];
        if (GRAD) Fx[k] = c1 * Fx[k];
      }
    } else if (gamma == 1.0) {
      acrn_fold<NO, EXACT, GRAD>(x + L, u2, iu2, 0.0, 1.0, F, Fx);
      acrn_fold<NO, EXACT, GRAD>(x - L, u2, iu2, 0.0, 1.0, F, Fx);
      acrn_fold<NO, EXACT, GRAD>(x, u2, iu2, 0.0, -2.0, F, Fx);
#pragma unroll
      for (int k = 0; k < NO; ++k) {
        F[k] = c2 * F[k];
        if (GRAD) Fx[k] = c2 * Fx[k];
      }
    } else {
      double G[NO], Gx[NO];
#pragma unroll
      for (int k = 0; k < NO; ++k) {
        G[k] = 0.0;
        Gx[k] = 0.0;
      }
      acrn_fold<NO, EXACT, GRAD>(x + L, u2, iu2, 1.0, 0.0, F, Fx);
      acrn_fold<NO, EXACT, GRAD>(x - L, u2, iu2, -1.0, 0.0, F, Fx);
      acrn_fold<NO, EXACT, GRAD>(x + L, u2, iu2, 0.0, 1.0, G, Gx);
      acrn_fold<NO, EXACT, GRAD>(x - L, u2, iu2, 0.0, 1.0, G, Gx);
      acrn_fold<NO, EXACT, GRAD>(x, u2, iu2, 0.0, -2.0, G, Gx);
#pragma unroll
      for (int k = 0; k < NO; ++k) {
        F[k] = c1 * F[k] + c2 * G[k];
        if (GRAD) Fx[k] = c1 * Fx[k] + c2 * Gx[k];
      }
    }
  } else {
    const double iR2 = 1.0 / (x * x + u2);
    double f = M * sqrt(iR2);
#pragma unroll
    for (int k = 0; k < NO; ++k) {
      F[k] = f;
      f = f * iR2;
      if (GRAD) Fx[k] = -(2.0 * k + 1.0) * x * f;
    }
  }
}

// dehnen_bar.py:1318-1326 (without the -G factor)
__device__ __forceinline__ double t3_ph(const smx_potential& p, double x, double y, double z) {
  const double a = p.t3_a, b = p.t3_b;
  const double ze = sqrt(z * z + b * b);
  const double Z = ze + a;
  const double aZ = a * Z;
  double F[3];
  needle<3, true, false>(p.t3_M, p.t3_L, p.t3_gamma, x, y * y + Z * Z, F, nullptr);
  double ph = F[0];
  ph = ph + (a * (Z - a / 3.0)) * F[1];
  ph = ph + (aZ * aZ) * F[2];
  return ph;
}

// dehnen_bar.py:1371-1383 (without the -G factor)
__device__ __forceinline__ double v4_ph(const smx_potential& p, double x, double y, double z) {
  const double a = p.v4_a, b = p.v4_b;
  const double ze = sqrt(z * z + b * b);
  const double Z = ze + a;
  const double aZ = a * Z;
  const double hB2 = 0.5 * b * b;
  double F[5];
  needle<5, true, false>(p.v4_M, p.v4_L, p.v4_gamma, x, y * y + Z * Z, F, nullptr);
  double ph = F[0];
  ph = ph + (a * (Z - 0.4 * a) + hB2) * F[1];
  ph = ph + ((3.0 * a) * ((0.2 * aZ) * ((Z + Z) - a) + hB2 * (Z - 0.2 * a))) * F[2];
  ph = ph + ((aZ * aZ) * (aZ + 6.0 * hB2)) * F[3];
  ph = ph + (((((7.0 * hB2) * ((aZ * aZ) * aZ)) * Z) / ze)) * F[4];
  return ph;
}

// likelihood.py:46-71: NFW + MN + T3 + V4
__device__ __forceinline__ double phi_total(const smx_potential& p, double x, double y, double z) {
  const double h = nfw_phi(p, x, y, z);
  const double m = mn_phi(p, x, y, z);
  const double t = (-p.G) * t3_ph(p, x, y, z);
  const double v = (-p.G) * v4_ph(p, x, y, z);
  return ((h + m) + t) + v;
}

// ---------------------------------------------------------------------------
// Fast force: the same derivative jax.grad takes of the potential
// (model.py:152-157), formed analytically instead of by dual numbers.
//
// For a needle kernel F_n(x, u2) = int K(x') M ((x - x')^2 + u2)^(-n/2) dx',
//   dF_n/dx  = c1 (Ax_n(x+L) - Ax_n(x-L)) + c2 (A_n(x+L) + A_n(x-L) - 2 A_n(x))
//              (dA_n/dx = Ax_n = r^-n, dC_n/dx = A_n),
//   dF_n/du2 = -(n/2) F_{n+2},
// so the recurrence run two orders further gives the whole gradient in plain
// doubles.  NFW and Miyamoto-Nagai use their closed-form derivatives.
// ---------------------------------------------------------------------------

// grad of T3's bracket ph (dehnen_bar.py:1318-1326): Phi_T3 = -G ph
__device__ __forceinline__ void t3_grad(const smx_potential& p, double x, double y, double z,
                                        double& gx, double& gy, double& gz) {
  const double a = p.t3_a, b = p.t3_b;
  const double ze = sqrt(z * z + b * b);
  const double Z = ze + a;
  const double u2 = y * y + Z * Z;
  double F[4], Fx[4];
  needle<4, false, true>(p.t3_M, p.t3_L, p.t3_gamma, x, u2, F, Fx);
  const double c[3] = {1.0, a * (Z - a / 3.0), (a * Z) * (a * Z)};
  const double dc[3] = {0.0, a, 2.0 * a * a * Z};
  double sx = 0.0, su = 0.0, sz = 0.0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    sx += c[k] * Fx[k];
    su += c[k] * (-(k + 0.5) * F[k + 1]);
    sz += dc[k] * F[k];
  }
  gx = sx;
  gy = su * 2.0 * y;
  gz = (sz + su * 2.0 * Z) * (z / ze);
}

// grad of V4's bracket ph (dehnen_bar.py:1371-1383): Phi_V4 = -G ph
__device__ __forceinline__ void v4_grad(const smx_potential& p, double x, double y, double z,
                                        double& gx, double& gy, double& gz) {
  const double a = p.v4_a, b = p.v4_b;
  const double ze = sqrt(z * z + b * b);
  const double Z = ze + a;
  const double u2 = y * y + Z * Z;
  const double hB2 = 0.5 * b * b;
  double F[6], Fx[6];
  needle<6, false, true>(p.v4_M, p.v4_L, p.v4_gamma, x, u2, F, Fx);
  const double a2 = a * a, a3 = a2 * a, Z2 = Z * Z;
  const double ize = 1.0 / ze;
  const double c[5] = {1.0, a * (Z - 0.4 * a) + hB2,
                       3.0 * a * (0.2 * a * Z * (2.0 * Z - a) + hB2 * (Z - 0.2 * a)),
                       a3 * Z2 * Z + 6.0 * hB2 * a2 * Z2, 7.0 * hB2 * a3 * Z2 * Z2 * ize};
  // d/dZ with ze = Z - a
  const double dc[5] = {0.0, a, 3.0 * a * (0.2 * a * (4.0 * Z - a) + hB2),
                        3.0 * a3 * Z2 + 12.0 * hB2 * a2 * Z,
                        7.0 * hB2 * a3 * Z2 * Z * (3.0 * Z - 4.0 * a) * ize * ize};
  double sx = 0.0, su = 0.0, sz = 0.0;
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    sx += c[k] * Fx[k];
    su += c[k] * (-(k + 0.5) * F[k + 1]);
    sz += dc[k] * F[k];
  }
  gx = sx;
  gy = su * 2.0 * y;
  gz = (sz + su * 2.0 * Z) * (z / ze);
}

// -grad Phi of the canonical potential with closed-form derivatives.
__device__ __forceinline__ void accel_fast(const smx_potential& p, double x, double y, double z,
                                           double& ax, double& ay, double& az) {
  // NFW (potentials.py:18-27): Phi = mGM log(1 + r/Rs) / (r + eps), r = sqrt(q.q + eps)
  double gx, gy, gz;
  {
    const double* R = p.nfw_rot;
    const double dx = x - p.nfw_origin[0], dy = y - p.nfw_origin[1], dz = z - p.nfw_origin[2];
    const double rx = (R[0] * dx + R[1] * dy) + R[2] * dz;
    const double ry = (R[3] * dx + R[4] * dy) + R[5] * dz;
    const double rz = (R[6] * dx + R[7] * dy) + R[8] * dz;
    const double qx = rx / p.nfw_a, qy = ry / p.nfw_b, qz = rz / p.nfw_c;
    const double r = sqrt(((qx * qx + qy * qy) + qz * qz) + kEps);
    const double re = r + kEps;
    const double dPhi_dr = p.nfw_mGM * ((1.0 / (p.nfw_Rs + r)) / re - log(1.0 + r / p.nfw_Rs) / (re * re));
    const double f = dPhi_dr / r;
    const double gqx = f * qx / p.nfw_a, gqy = f * qy / p.nfw_b, gqz = f * qz / p.nfw_c;
    gx = R[0] * gqx + R[3] * gqy + R[6] * gqz;
    gy = R[1] * gqx + R[4] * gqy + R[7] * gqz;
    gz = R[2] * gqx + R[5] * gqy + R[8] * gqz;
  }
  // Miyamoto-Nagai (potentials.py:51-65): Phi = mGM / sqrt(R^2 + (Rs + zeta)^2)
  {
    const double* R = p.mn_rot;
    const double dx = x - p.mn_origin[0], dy = y - p.mn_origin[1], dz = z - p.mn_origin[2];
    const double rx = ((R[0] * dx + R[1] * dy) + R[2] * dz) + kEps;
    const double ry = ((R[3] * dx + R[4] * dy) + R[5] * dz) + kEps;
    const double rz = ((R[6] * dx + R[7] * dy) + R[8] * dz) + kEps;
    const double zeta = sqrt(rz * rz + p.mn_Hs * p.mn_Hs);
    const double u = p.mn_Rs + zeta;
    const double d2 = (rx * rx + ry * ry) + u * u;
    const double id = 1.0 / sqrt(d2);
    const double f = -p.mn_mGM * id * id * id;
    const double grx = f * rx, gry = f * ry, grz = f * u * (rz / zeta);
    gx += R[0] * grx + R[3] * gry + R[6] * grz;
    gy += R[1] * grx + R[4] * gry + R[7] * grz;
    gz += R[2] * grx + R[5] * gry + R[8] * grz;
  }
  double tx, ty, tz, vx, vy, vz;
  t3_grad(p, x, y, z, tx, ty, tz);
  v4_grad(p, x, y, z, vx, vy, vz);
  // a = -grad Phi, Phi_bar = -G ph
  ax = -gx + p.G * (tx + vx);
  ay = -gy + p.G * (ty + vy);
  az = -gz + p.G * (tz + vz);
}

// dehnen_bar.py:1425-1438
__device__ __forceinline__ double t3_rho(const smx_potential& p, double x, double y, double z) {
  const double a = p.t3_a, b = p.t3_b;
  const double ze = sqrt(z * z + b * b);
  const double iz = 1.0 / ze;
  const double Z = ze + a;
  const double Z2 = Z * Z, A2 = a * a;
  double F[5];
  needle<5, true, false>(p.t3_M, p.t3_L, p.t3_gamma, x, y * y + Z * Z, F, nullptr);
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
  needle<7, true, false>(p.v4_M, p.v4_L, p.v4_gamma, x, y * y + Z * Z, F, nullptr);
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
