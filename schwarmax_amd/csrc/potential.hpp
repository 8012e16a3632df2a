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
#include "dual.hpp"

namespace smx {

constexpr double kEps = 1e-12;   // constants.py:11
constexpr double kPi = 3.141592653589793;

// Every function is a template on the scalar type S (double, or Dual<N> for
// derivatives); parameters that can carry a tangent are passed as S, the rest
// are read from smx_potential.  With S = double the arithmetic is the value
// path, operation for operation.

template <class S>
__device__ __forceinline__ void rotate3(const double* R, const double* o, const S& x, const S& y,
                                        const S& z, S& rx, S& ry, S& rz) {
  const S dx = x - o[0], dy = y - o[1], dz = z - o[2];
  rx = (R[0] * dx + R[1] * dy) + R[2] * dz;
  ry = (R[3] * dx + R[4] * dy) + R[5] * dz;
  rz = (R[6] * dx + R[7] * dy) + R[8] * dz;
}

// potentials.py:18-27
template <class S>
__device__ __forceinline__ S nfw_phi_t(const smx_potential& p, const S& mGM, const S& Rs, const S& x,
                                       const S& y, const S& z) {
  S rx, ry, rz;
  rotate3(p.nfw_rot, p.nfw_origin, x, y, z, rx, ry, rz);
  const S qx = rx / p.nfw_a, qy = ry / p.nfw_b, qz = rz / p.nfw_c;
  const S r = sqrt(((qx * qx + qy * qy) + qz * qz) + kEps);
  return (mGM * log(1.0 + r / Rs)) / (r + kEps);
}

// potentials.py:51-65
template <class S>
__device__ __forceinline__ S mn_phi_t(const smx_potential& p, const S& mGM, const S& Rs, const S& Hs,
                                      const S& x, const S& y, const S& z) {
  S rx, ry, rz;
  rotate3(p.mn_rot, p.mn_origin, x, y, z, rx, ry, rz);
  rx = rx + kEps; ry = ry + kEps; rz = rz + kEps;
  const S R = sqrt(rx * rx + ry * ry);
  const S u = Rs + sqrt(rz * rz + Hs * Hs);
  const S d2 = R * R + u * u;
  return mGM / sqrt(d2);
}

// potentials.py:29-49
template <class S>
__device__ __forceinline__ S mn_rho_t(const smx_potential& p, const S& rho0, const S& Rs, const S& Hs,
                                      const S& x, const S& y, const S& z) {
  S rx, ry, rz;
  rotate3(p.mn_rot, p.mn_origin, x, y, z, rx, ry, rz);
  rx += kEps; ry += kEps; rz += kEps;
  const S R = sqrt(rx * rx + ry * ry);
  const S beta = sqrt(rz * rz + Hs * Hs);
  const S sb = Rs + beta;
  const S D2 = R * R + sb * sb;
  const S num = Rs * R * R + (Rs + 3.0 * beta) * (sb * sb);
  const S den = (beta * beta * beta) * pow(D2, 2.5);
  return rho0 * (num / den);
}

// Run the reference's recurrence (dehnen_bar.py:104-157, odd branch; C_n =
// x A_n - B_n, 269-278) at abscissa xp and fold each order into running sums:
//   F[k] += sA * A_k + sC * C_k        (and, with the gradient, Q[k] += sA * Ax_k + sC * A_k)
// sA, sC in {0, +-1, -2}: the sums reproduce the reference's differences
// (A(x+L) - A(x-L)) and ((C(x+L) + C(x-L)) - 2 C(x)) operation for operation.
// EXACT divides by cur*u2 as the reference does; otherwise multiplies by a
// precomputed 1/u2 (force path).
template <int NO, bool EXACT, bool GRAD, class S>
__device__ __forceinline__ void acrn_fold(const S& xp, const S& u2, const S& iu2, double sA, double sC,
                                          S* F, S* Q) {
  const S R2 = xp * xp + u2;
  const S r = sqrt(R2);
  S ax = 1.0 / r;
  // 1/R2: a divide on the value path, the square of 1/r on the force path
  const S iR2 = EXACT ? 1.0 / R2 : ax * ax;
  S a = log(xp + r);
  S b = r;
#pragma unroll
  for (int k = 0; k < NO; ++k) {
    if (k > 0) {
      const double cur = 2.0 * k - 1.0;
      const S num = (cur - 1.0) * a + xp * ax;
      if (EXACT) {
        a = num / (cur * u2);
        b = (-ax) / cur;
      } else {
        // force path: divisions by the unrolled order's constant become
        // products with its reciprocal (one rounding more, no FP64 divide)
        constexpr double one = 1.0;
        const double icur = one / cur;
        a = num * (iu2 * icur);
        b = (-ax) * icur;
      }
      ax = ax * iR2;
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
// GRAD also dF/dx.  SL is the type of the needle length (double when it
// carries no tangent).
template <int NO, bool EXACT, bool GRAD, class S, class SL>
__device__ __forceinline__ void needle(const S& M, const SL& L, double gamma, const S& x, const S& u2,
                                       S* F, S* Fx) {
#pragma unroll
  for (int k = 0; k < NO; ++k) {
    F[k] = Lift<S>::c(0.0);
    if (GRAD) Fx[k] = Lift<S>::c(0.0);
  }
  if (val(L) > 0.0) {
    const S iu2 = EXACT ? Lift<S>::c(0.0) : 1.0 / u2;
    if (gamma == 0.0) {
      const S c1 = M * (1.0 - gamma) / (2.0 * L);
      acrn_fold<NO, EXACT, GRAD>(x + L, u2, iu2, 1.0, 0.0, F, Fx);
      acrn_fold<NO, EXACT, GRAD>(x - L, u2, iu2, -1.0, 0.0, F, Fx);
#pragma unroll
      for (int k = 0; k < NO; ++k) {
        F[k] = c1 * F[k];
        if (GRAD) Fx[k] = c1 * Fx[k];
      }
    } else if (gamma == 1.0) {
      const S c2 = M * gamma / (L * L);
      acrn_fold<NO, EXACT, GRAD>(x + L, u2, iu2, 0.0, 1.0, F, Fx);
      acrn_fold<NO, EXACT, GRAD>(x - L, u2, iu2, 0.0, 1.0, F, Fx);
      acrn_fold<NO, EXACT, GRAD>(x, u2, iu2, 0.0, -2.0, F, Fx);
#pragma unroll
      for (int k = 0; k < NO; ++k) {
        F[k] = c2 * F[k];
        if (GRAD) Fx[k] = c2 * Fx[k];
      }
    } else {
      const S c1 = M * (1.0 - gamma) / (2.0 * L);
      const S c2 = M * gamma / (L * L);
      S G[NO], Gx[NO];
#pragma unroll
      for (int k = 0; k < NO; ++k) {
        G[k] = Lift<S>::c(0.0);
        Gx[k] = Lift<S>::c(0.0);
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
    const S iR2 = 1.0 / (x * x + u2);
    S f = M * sqrt(iR2);
#pragma unroll
    for (int k = 0; k < NO; ++k) {
      F[k] = f;
      f = f * iR2;
      if (GRAD) Fx[k] = (-(2.0 * k + 1.0)) * x * f;
    }
  }
}

// dehnen_bar.py:1318-1326 (without the -G factor)
template <class S>
__device__ __forceinline__ S t3_ph_t(const smx_potential& p, const S& M, const S& L, const S& a,
                                     const S& b, const S& x, const S& y, const S& z) {
  const S ze = sqrt(z * z + b * b);
  const S Z = ze + a;
  const S aZ = a * Z;
  S F[3];
  needle<3, true, false>(M, L, p.t3_gamma, x, y * y + Z * Z, F, (S*)nullptr);
  S ph = F[0];
  ph = ph + (a * (Z - a / 3.0)) * F[1];
  ph = ph + (aZ * aZ) * F[2];
  return ph;
}

// dehnen_bar.py:1371-1383 (without the -G factor)
template <class S>
__device__ __forceinline__ S v4_ph_t(const smx_potential& p, const S& M, const S& x, const S& y,
                                     const S& z) {
  const double a = p.v4_a, b = p.v4_b;
  const S ze = sqrt(z * z + b * b);
  const S Z = ze + a;
  const S aZ = a * Z;
  const double hB2 = 0.5 * b * b;
  S F[5];
  needle<5, true, false>(M, p.v4_L, p.v4_gamma, x, y * y + Z * Z, F, (S*)nullptr);
  S ph = F[0];
  ph = ph + (a * (Z - 0.4 * a) + hB2) * F[1];
  ph = ph + ((3.0 * a) * ((0.2 * aZ) * ((Z + Z) - a) + hB2 * (Z - 0.2 * a))) * F[2];
  ph = ph + ((aZ * aZ) * (aZ + 6.0 * hB2)) * F[3];
  ph = ph + (((((7.0 * hB2) * ((aZ * aZ) * aZ)) * Z) / ze)) * F[4];
  return ph;
}

// ---------------------------------------------------------------------------
// Gradients: the same derivative jax.grad takes of the potential
// (model.py:152-157), formed analytically.
//
// For a needle kernel F_n(x, u2) = int K(x') M ((x - x')^2 + u2)^(-n/2) dx',
//   dF_n/dx  = c1 (Ax_n(x+L) - Ax_n(x-L)) + c2 (A_n(x+L) + A_n(x-L) - 2 A_n(x))
//              (dA_n/dx = Ax_n = r^-n, dC_n/dx = A_n),
//   dF_n/du2 = -(n/2) F_{n+2},
// so the recurrence run two orders further gives the whole gradient.  NFW and
// Miyamoto-Nagai use their closed-form derivatives.  Each returns +grad Phi
// of its component (for the bar components, of the bracket ph: Phi = -G ph).
// ---------------------------------------------------------------------------

// NFW (potentials.py:18-27): Phi = mGM log(1 + r/Rs) / (r + eps), r = sqrt(q.q + eps)
template <class S>
__device__ __forceinline__ void nfw_grad_t(const smx_potential& p, const S& mGM, const S& Rs,
                                           const S& x, const S& y, const S& z, S& gx, S& gy, S& gz) {
  const double* R = p.nfw_rot;
  const S dx = x - p.nfw_origin[0], dy = y - p.nfw_origin[1], dz = z - p.nfw_origin[2];
  const S rx = (R[0] * dx + R[1] * dy) + R[2] * dz;
  const S ry = (R[3] * dx + R[4] * dy) + R[5] * dz;
  const S rz = (R[6] * dx + R[7] * dy) + R[8] * dz;
  // products with the axis reciprocals (invariant across calls) instead of
  // FP64 divides; one rounding more than the quotients
  const double ia = 1.0 / p.nfw_a, ib = 1.0 / p.nfw_b, ic = 1.0 / p.nfw_c;
  const S qx = rx * ia, qy = ry * ib, qz = rz * ic;
  const S r = sqrt(((qx * qx + qy * qy) + qz * qz) + kEps);
  const S ire = 1.0 / (r + kEps);
  const S dPhi_dr = mGM * ((1.0 / (Rs + r)) * ire - log(1.0 + r / Rs) * (ire * ire));
  const S f = dPhi_dr / r;
  const S gqx = (f * qx) * ia, gqy = (f * qy) * ib, gqz = (f * qz) * ic;
  gx = (R[0] * gqx + R[3] * gqy) + R[6] * gqz;
  gy = (R[1] * gqx + R[4] * gqy) + R[7] * gqz;
  gz = (R[2] * gqx + R[5] * gqy) + R[8] * gqz;
}

// Miyamoto-Nagai (potentials.py:51-65): Phi = mGM / sqrt(R^2 + (Rs + zeta)^2)
template <class S>
__device__ __forceinline__ void mn_grad_t(const smx_potential& p, const S& mGM, const S& Rs,
                                          const S& Hs, const S& x, const S& y, const S& z, S& gx,
                                          S& gy, S& gz) {
  const double* R = p.mn_rot;
  const S dx = x - p.mn_origin[0], dy = y - p.mn_origin[1], dz = z - p.mn_origin[2];
  const S rx = ((R[0] * dx + R[1] * dy) + R[2] * dz) + kEps;
  const S ry = ((R[3] * dx + R[4] * dy) + R[5] * dz) + kEps;
  const S rz = ((R[6] * dx + R[7] * dy) + R[8] * dz) + kEps;
  const S zeta = sqrt(rz * rz + Hs * Hs);
  const S u = Rs + zeta;
  const S d2 = (rx * rx + ry * ry) + u * u;
  const S id = 1.0 / sqrt(d2);
  const S f = -mGM * id * id * id;
  const S grx = f * rx, gry = f * ry, grz = f * u * (rz * (1.0 / zeta));
  gx = (R[0] * grx + R[3] * gry) + R[6] * grz;
  gy = (R[1] * grx + R[4] * gry) + R[7] * grz;
  gz = (R[2] * grx + R[5] * gry) + R[8] * grz;
}

// grad of T3's bracket ph (dehnen_bar.py:1318-1326): Phi_T3 = -G ph
template <class S>
__device__ __forceinline__ void t3_grad_t(const smx_potential& p, const S& M, const S& L, const S& a,
                                          const S& b, const S& x, const S& y, const S& z, S& gx,
                                          S& gy, S& gz) {
  const S ze = sqrt(z * z + b * b);
  const S Z = ze + a;
  const S u2 = y * y + Z * Z;
  S F[4], Fx[4];
  needle<4, false, true>(M, L, p.t3_gamma, x, u2, F, Fx);
  const S c1 = a * (Z - a / 3.0), c2 = (a * Z) * (a * Z);
  const S dc2 = 2.0 * a * a * Z;
  S sx = Fx[0] + c1 * Fx[1] + c2 * Fx[2];
  S su = (-0.5) * F[1] + c1 * ((-1.5) * F[2]) + c2 * ((-2.5) * F[3]);
  S sz = a * F[1] + dc2 * F[2];
  gx = sx;
  gy = su * 2.0 * y;
  gz = (sz + su * 2.0 * Z) * (z * (1.0 / ze));
}

// grad of V4's bracket ph (dehnen_bar.py:1371-1383): Phi_V4 = -G ph
template <class S>
__device__ __forceinline__ void v4_grad_t(const smx_potential& p, const S& M, const S& x, const S& y,
                                          const S& z, S& gx, S& gy, S& gz) {
  const double a = p.v4_a, b = p.v4_b;
  const S ze = sqrt(z * z + b * b);
  const S Z = ze + a;
  const S u2 = y * y + Z * Z;
  const double hB2 = 0.5 * b * b;
  S F[6], Fx[6];
  needle<6, false, true>(M, p.v4_L, p.v4_gamma, x, u2, F, Fx);
  const double a2 = a * a, a3 = a2 * a;
  const S Z2 = Z * Z;
  const S ize = 1.0 / ze;
  const S c[5] = {Lift<S>::c(1.0), a * (Z - 0.4 * a) + hB2,
                  3.0 * a * (0.2 * a * Z * (2.0 * Z - a) + hB2 * (Z - 0.2 * a)),
                  a3 * Z2 * Z + 6.0 * hB2 * a2 * Z2, 7.0 * hB2 * a3 * Z2 * Z2 * ize};
  // d/dZ with ze = Z - a
  const S dc[5] = {Lift<S>::c(0.0), Lift<S>::c(a), 3.0 * a * (0.2 * a * (4.0 * Z - a) + hB2),
                   3.0 * a3 * Z2 + 12.0 * hB2 * a2 * Z,
                   7.0 * hB2 * a3 * Z2 * Z * (3.0 * Z - 4.0 * a) * ize * ize};
  S sx = Lift<S>::c(0.0), su = Lift<S>::c(0.0), sz = Lift<S>::c(0.0);
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    sx += c[k] * Fx[k];
    su += c[k] * ((-(k + 0.5)) * F[k + 1]);
    sz += dc[k] * F[k];
  }
  gx = sx;
  gy = su * 2.0 * y;
  gz = (sz + su * 2.0 * Z) * (z * ize);
}

// dehnen_bar.py:1425-1438
template <class S>
__device__ __forceinline__ S t3_rho_t(const smx_potential& p, const S& M, const S& L, const S& a,
                                      const S& b, const S& x, const S& y, const S& z) {
  const S ze = sqrt(z * z + b * b);
  const S iz = 1.0 / ze;
  const S Z = ze + a;
  const S Z2 = Z * Z, A2 = a * a;
  S F[5];
  needle<5, true, false>(M, L, p.t3_gamma, x, y * y + Z * Z, F, (S*)nullptr);
  S rh = (3.0 * ((ze * ze) * ze)) * F[2];
  rh = rh + (((5.0 * a) * Z2) * (((3.0 * ze) - (2.0 * a)) * ze + A2)) * F[3];
  rh = rh + ((((35.0 * A2) * Z2) * Z2) * ze) * F[4];
  return (((b * b) * ((iz * iz) * iz)) * (0.25 / kPi)) * rh;
}

// dehnen_bar.py:1526-1543
template <class S>
__device__ __forceinline__ S v4_rho_t(const smx_potential& p, const S& M, const S& x, const S& y,
                                      const S& z) {
  const double a = p.v4_a, b = p.v4_b;
  const S ze = sqrt(z * z + b * b);
  const S iz = 1.0 / ze;
  const S Z = ze + a;
  const double A2 = a * a;
  const S Z2 = Z * Z, AZ = a * Z;
  S F[7];
  needle<7, true, false>(M, p.v4_L, p.v4_gamma, x, y * y + Z * Z, F, (S*)nullptr);
  S rh = (5.0 * ((ze * ze) * ze)) * F[3];
  rh = rh + (((7.0 * a) * Z2) * (((((5.0 * ze) * ze - (4.0 * a) * ze) + 3.0 * A2) -
                                  ((2.0 * a) * A2) * iz) + ((A2 * A2) * iz) * iz)) * F[4];
  rh = rh + ((((63.0 * A2) * Z2) * Z2) * (((ze + ze) - 2.0 * a) + A2 * iz)) * F[5];
  rh = rh + ((((((231.0 * A2) * AZ) * Z2) * Z2) * Z)) * F[6];
  return (((b * b) * ((iz * iz) * iz)) * (((0.375 * b) * b) / kPi)) * rh;
}

// ---------------------------------------------------------------------------
// The potential's parameters that depend on theta, as scalars of type S
// (derivatives of the ICs and the density target are taken with respect to
// these fields; the map theta -> fields is differentiated on the host).
// ---------------------------------------------------------------------------
enum PotField {
  PF_NFW_MGM = 0, PF_NFW_RS, PF_MN_MGM, PF_MN_RHO0, PF_MN_RS, PF_MN_HS, PF_T3_M, PF_T3_L, PF_T3_A,
  PF_T3_B, PF_V4_M, PF_COUNT
};

template <class S>
struct PotFields {
  S nfw_mGM, nfw_Rs, mn_mGM, mn_rho0, mn_Rs, mn_Hs, t3_M, t3_L, t3_a, t3_b, v4_M;
};

__device__ __forceinline__ double field_value(const smx_potential& p, int f) {
  switch (f) {
    case PF_NFW_MGM: return p.nfw_mGM;
    case PF_NFW_RS: return p.nfw_Rs;
    case PF_MN_MGM: return p.mn_mGM;
    case PF_MN_RHO0: return p.mn_rho0;
    case PF_MN_RS: return p.mn_Rs;
    case PF_MN_HS: return p.mn_Hs;
    case PF_T3_M: return p.t3_M;
    case PF_T3_L: return p.t3_L;
    case PF_T3_A: return p.t3_a;
    case PF_T3_B: return p.t3_b;
    default: return p.v4_M;
  }
}

// fields as constants (S = double: the values themselves)
template <class S>
__device__ __forceinline__ PotFields<S> fields_const(const smx_potential& p) {
  PotFields<S> f;
  f.nfw_mGM = Lift<S>::c(p.nfw_mGM); f.nfw_Rs = Lift<S>::c(p.nfw_Rs);
  f.mn_mGM = Lift<S>::c(p.mn_mGM); f.mn_rho0 = Lift<S>::c(p.mn_rho0);
  f.mn_Rs = Lift<S>::c(p.mn_Rs); f.mn_Hs = Lift<S>::c(p.mn_Hs);
  f.t3_M = Lift<S>::c(p.t3_M); f.t3_L = Lift<S>::c(p.t3_L);
  f.t3_a = Lift<S>::c(p.t3_a); f.t3_b = Lift<S>::c(p.t3_b);
  f.v4_M = Lift<S>::c(p.v4_M);
  return f;
}

// fields with tangent slot k seeded on field f0 + k, k < N
template <int N>
__device__ __forceinline__ PotFields<Dual<N>> fields_seeded(const smx_potential& p, int f0) {
  Dual<N> v[PF_COUNT];
#pragma unroll
  for (int f = 0; f < PF_COUNT; ++f) {
    v[f] = dconst<N>(field_value(p, f));
#pragma unroll
    for (int k = 0; k < N; ++k)
      if (f == f0 + k) v[f].d[k] = 1.0;
  }
  PotFields<Dual<N>> r;
  r.nfw_mGM = v[PF_NFW_MGM]; r.nfw_Rs = v[PF_NFW_RS]; r.mn_mGM = v[PF_MN_MGM];
  r.mn_rho0 = v[PF_MN_RHO0]; r.mn_Rs = v[PF_MN_RS]; r.mn_Hs = v[PF_MN_HS];
  r.t3_M = v[PF_T3_M]; r.t3_L = v[PF_T3_L]; r.t3_a = v[PF_T3_A]; r.t3_b = v[PF_T3_B];
  r.v4_M = v[PF_V4_M];
  return r;
}

// likelihood.py:46-71: NFW + MN + T3 + V4
template <class S>
__device__ __forceinline__ S phi_total_t(const smx_potential& p, const PotFields<S>& f, const S& x,
                                         const S& y, const S& z) {
  const S h = nfw_phi_t(p, f.nfw_mGM, f.nfw_Rs, x, y, z);
  const S m = mn_phi_t(p, f.mn_mGM, f.mn_Rs, f.mn_Hs, x, y, z);
  const S t = (-p.G) * t3_ph_t(p, f.t3_M, f.t3_L, f.t3_a, f.t3_b, x, y, z);
  const S v = (-p.G) * v4_ph_t(p, f.v4_M, x, y, z);
  return ((h + m) + t) + v;
}

// likelihood.py:13-43: MN + T3 + V4
template <class S>
__device__ __forceinline__ S rho_total_t(const smx_potential& p, const PotFields<S>& f, const S& x,
                                         const S& y, const S& z) {
  return (mn_rho_t(p, f.mn_rho0, f.mn_Rs, f.mn_Hs, x, y, z) +
          t3_rho_t(p, f.t3_M, f.t3_L, f.t3_a, f.t3_b, x, y, z)) +
         v4_rho_t(p, f.v4_M, x, y, z);
}

// -grad Phi with all fields constant
template <class S>
__device__ __forceinline__ void accel_t(const smx_potential& p, const PotFields<S>& f, const S& x,
                                        const S& y, const S& z, S& ax, S& ay, S& az) {
  S gx, gy, gz, hx, hy, hz, tx, ty, tz, vx, vy, vz;
  nfw_grad_t(p, f.nfw_mGM, f.nfw_Rs, x, y, z, gx, gy, gz);
  mn_grad_t(p, f.mn_mGM, f.mn_Rs, f.mn_Hs, x, y, z, hx, hy, hz);
  t3_grad_t(p, f.t3_M, f.t3_L, f.t3_a, f.t3_b, x, y, z, tx, ty, tz);
  v4_grad_t(p, f.v4_M, x, y, z, vx, vy, vz);
  // a = -grad Phi, Phi_bar = -G ph
  ax = -(gx + hx) + p.G * (tx + vx);
  ay = -(gy + hy) + p.G * (ty + vy);
  az = -(gz + hz) + p.G * (tz + vz);
}

// ---------------------------------------------------------------------------
// double-valued entry points used by the forward kernels
// ---------------------------------------------------------------------------
__device__ __forceinline__ double phi_total(const smx_potential& p, double x, double y, double z) {
  return phi_total_t<double>(p, fields_const<double>(p), x, y, z);
}

__device__ __forceinline__ double rho_total(const smx_potential& p, double x, double y, double z) {
  return rho_total_t<double>(p, fields_const<double>(p), x, y, z);
}

__device__ __forceinline__ void accel_fast(const smx_potential& p, double x, double y, double z,
                                           double& ax, double& ay, double& az) {
  accel_t<double>(p, fields_const<double>(p), x, y, z, ax, ay, az);
}

}  // namespace smx
