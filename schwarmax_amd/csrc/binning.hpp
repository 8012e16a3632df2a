// Grid assignment and Gauss-Hermite terms for one phase-space sample.
//
// Index arithmetic follows utils.py:10-42 operation by operation with FMA
// contraction disabled, so cell and aperture ids are bit-identical to an IEEE
// evaluation of q = (p - min) / (max - min); i = clip(floor(q * n), 0, n - 1);
// out of bounds when q >= 1 or q < 0.
#pragma once
#include "common.hpp"
#include "../../include/smx.h"

namespace smx {

constexpr double kKpcGyrToKms = 0.9777922216807892;   // constants.py:10
constexpr double kInvSqrt2Pi = 0.3989422804014327;     // (2 pi)^-1/2, ghmoments.py:8
constexpr double kSq2 = 1.4142135623730951;
constexpr double kSq3 = 1.7320508075688772;

// One coordinate of assign_regular_grid.  Returns false when out of bounds.
__device__ __forceinline__ bool grid_coord(double p, double lo, double hi, int n, int& c) {
#pragma clang fp contract(off)
  const double q = (p - lo) / (hi - lo);
  const double f = floor(q * static_cast<double>(n));
  int ci = (f != f) ? 0 : (f < 0.0 ? 0 : (f > static_cast<double>(n - 1) ? n - 1 : static_cast<int>(f)));
  c = ci;
  return !(q >= 1.0 || q < 0.0);
}

// (R, z, phi) cell id or -1 (integrates.py:121-129).
__device__ __forceinline__ int rzphi_cell(const smx_grids& g, double R, double z, double phi) {
  int a, b, c;
  const bool ok0 = grid_coord(R, g.rz_min[0], g.rz_max[0], g.nR, a);
  const bool ok1 = grid_coord(z, g.rz_min[1], g.rz_max[1], g.nz, b);
  const bool ok2 = grid_coord(phi, g.rz_min[2], g.rz_max[2], g.nphi, c);
  if (!(ok0 && ok1 && ok2)) return -1;
  return a + b * g.nR + c * g.nR * g.nz;
}

// Aperture id of sky position (xr, yr), or -1 when its map entry is not an
// aperture (integrates.py:132-143).  An off-field sample takes the pixel id
// nX*nY (utils.py:40-42), so it reads the map's last entry, bin_mapping[nX*nY]:
// the sentinel -1 when the map carries one; the last pixel's aperture when the
// host padded a map without one (the reference's gather clamps there).  The
// map therefore always has nX*nY + 1 entries.
__device__ __forceinline__ int aperture(const smx_grids& g, const int32_t* bin_mapping,
                                        double xr, double yr) {
  int a, b;
  const bool ok0 = grid_coord(xr, g.xy_min[0], g.xy_max[0], g.nX, a);
  const bool ok1 = grid_coord(yr, g.xy_min[1], g.xy_max[1], g.nY, b);
  const int pix = (ok0 && ok1) ? a + b * g.nX : g.nX * g.nY;
  const int v = bin_mapping[pix];
  return (v >= 0 && v < g.num_vbin) ? v : -1;
}

// Gauss weight and H0..H4 (ghmoments.py:6-36).
__device__ __forceinline__ void gh_terms(double vlos, double v0, double s, double& gw, double H[5]) {
#pragma clang fp contract(off)
  const double y = (vlos - v0) / s;
  gw = (kInvSqrt2Pi / s) * exp(-0.5 * (y * y));
  H[0] = 1.0;
  H[1] = kSq2 * y;
  H[2] = ((kSq2 * y) * H[1] - 1.0 * H[0]) / kSq2;
  H[3] = ((kSq2 * y) * H[2] - kSq2 * H[1]) / kSq3;
  H[4] = ((kSq2 * y) * H[3] - kSq3 * H[2]) / 2.0;
}

// The four symmetry copies of a sample (integrates.py:74-79, 117): copy c
// flips (x, y, vx, vy) when c >= 2 and (z, vz) when c is odd.  The products
// are formed once and the copies differ only by signs, which IEEE rounding
// preserves exactly, so each copy equals a literal evaluation on the flipped
// state.
struct SampleProj {
  double R, phi_p, phi_m;     // phi of (x, y) and of (-x, -y)
  double xa, xb, ya, yb;      // x' = sxy*xa + sz*xb ; y' = sxy*ya + sz*yb
  double va, vb;              // v_los = (sxy*va + sz*vb) * KPCGYR_TO_KMS
};

__device__ __forceinline__ void project(const smx_grids& g, const double* y, bool both_phi,
                                        SampleProj& s) {
#pragma clang fp contract(off)
  s.R = sqrt(y[0] * y[0] + y[1] * y[1]);
  s.phi_p = atan2(y[1], y[0]);
  s.phi_m = both_phi ? atan2(-y[1], -y[0]) : 0.0;
  const double* M = g.rot;
  s.xa = M[0] * y[0] + M[1] * y[1];
  s.xb = M[2] * y[2];
  s.ya = M[3] * y[0] + M[4] * y[1];
  s.yb = M[5] * y[2];
  s.va = M[6] * y[3] + M[7] * y[4];
  s.vb = M[8] * y[5];
}

__device__ __forceinline__ void copy_coords(const SampleProj& s, int c, double& z_sign, double& phi,
                                            double& xr, double& yr, double& vlos) {
#pragma clang fp contract(off)
  const double sxy = (c >= 2) ? -1.0 : 1.0;
  const double sz = (c & 1) ? -1.0 : 1.0;
  z_sign = sz;
  phi = (c >= 2) ? s.phi_m : s.phi_p;
  xr = sxy * s.xa + sz * s.xb;
  yr = sxy * s.ya + sz * s.yb;
  vlos = (sxy * s.va + sz * s.vb) * kKpcGyrToKms;
}

}  // namespace smx
