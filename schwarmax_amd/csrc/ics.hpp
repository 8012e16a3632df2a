// Jeans initial conditions of one star (jeans.py:9-94), its rotation-curve
// jitter scale and orbital time scale (model.py:117-169, utils.py:182-273),
// and the density target of one (R, z, phi) cell (model.py:192-213).
//
// Templates on the scalar type S: the forward kernels use double; the
// parameter Jacobians use Dual<N> with tangents seeded on potential fields.
// Positions, quadrature abscissae and finite-difference steps are constants.
#pragma once
#include "common.hpp"
#include "potential.hpp"

namespace smx {

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
template <int N>
__device__ __forceinline__ Dual<N> wave_sum(Dual<N> v) {
  v.v = wave_sum(v.v);
#pragma unroll
  for (int k = 0; k < N; ++k) v.d[k] = wave_sum(v.d[k]);
  return v;
}

// NaN-propagating max / min against a constant (jnp.maximum / jnp.minimum);
// the selected operand's tangent is kept
template <class S> __device__ __forceinline__ S smax(const S& a, double b) {
  const double v = val(a);
  return (v != v) ? a : (v > b ? a : Lift<S>::c(b));
}
template <class S> __device__ __forceinline__ S smin(const S& a, double b) {
  const double v = val(a);
  return (v != v) ? a : (v < b ? a : Lift<S>::c(b));
}

// point i of numpy's linspace(lo, hi, npts)
__device__ __forceinline__ double lin_pt(double lo, double hi, double step, int i, int npts) {
  return (i == npts - 1) ? hi : (static_cast<double>(i) * step + lo);
}

template <class S>
__device__ __forceinline__ S dphi_dz_t(const smx_potential& p, const PotFields<S>& f, double x,
                                       double y, double z) {
  const double d = 5e-3;
  const S X = Lift<S>::c(x), Y = Lift<S>::c(y);
  return (phi_total_t(p, f, X, Y, Lift<S>::c(z + d)) - phi_total_t(p, f, X, Y, Lift<S>::c(z - d))) /
         (2.0 * d);
}

// integral_{|z*|}^{10} rho(x, y, z') dPhi/dz(x, y, z') dz' by the trapezoid
// rule on npts points (jeans.py:37, 57), this lane's share; the caller sums
// over the wave.
template <class S>
__device__ S pressure_part_t(const smx_potential& p, const PotFields<S>& f, double x, double y,
                             double zlo, int npts, int lane) {
  const double hi = 10.0;
  const double step = (hi - zlo) / static_cast<double>(npts - 1);
  S acc = Lift<S>::c(0.0);
  for (int i = lane; i < npts; i += 64) {
    const double zi = lin_pt(zlo, hi, step, i, npts);
    const S fv = rho_total_t(p, f, Lift<S>::c(x), Lift<S>::c(y), Lift<S>::c(zi)) *
                 dphi_dz_t(p, f, x, y, zi);
    double w = 0.0;
    if (i > 0) w += zi - lin_pt(zlo, hi, step, i - 1, npts);
    if (i < npts - 1) w += lin_pt(zlo, hi, step, i + 1, npts) - zi;
    acc += (0.5 * w) * fv;
  }
  return acc;
}

template <class S>
__device__ __forceinline__ S phi_at(const smx_potential& p, const PotFields<S>& f, double x, double y,
                                    double z) {
  return phi_total_t(p, f, Lift<S>::c(x), Lift<S>::c(y), Lift<S>::c(z));
}

template <class S>
struct StarIC {
  S wn[6];      // Jeans position + velocity (before jitter)
  S vsc;        // velocity jitter scale clip(0.1 v_c, 1, 15)
  S T_orb;      // orbital time scale
};

constexpr int JEANS_NPTS = 1000;     // trapezoid points (jeans.py:37, 57)
constexpr double JEANS_DR = 5e-3;    // radial finite-difference step (jeans.py:60-61)

// The rest of star_ic, from the three pressure integrals at (x, y), (R + dR,
// 0) and (R - dR, 0).
template <class S>
__device__ void star_ic_finish(const smx_potential& p, const PotFields<S>& f, const double* w0,
                               const double* normals, int64_t i, const S& I0, const S& Ip,
                               const S& Im, StarIC<S>& o) {
  const double xs = w0[3 * i], ys = w0[3 * i + 1], zs = w0[3 * i + 2];
  const double Rs = sqrt(xs * xs + ys * ys);
  const double dR = JEANS_DR;
  const double b = 1.0;   // anisotropy_b (jeans.py:86)
  const S nu = rho_total_t(p, f, Lift<S>::c(xs), Lift<S>::c(ys), Lift<S>::c(zs));
  const S sz2 = smax((1.0 / nu) * I0, 0.0);
  const S sz = sqrt(sz2);
  const S sR2 = b * sz2;
  const S sR = sqrt(sR2);
  const S dnu = b * (Ip - Im) / (2.0 * dR);
  const double d = 5e-3;
  const S dphidR = (phi_at(p, f, Rs + d, 0.0, zs) - phi_at(p, f, Rs - d, 0.0, zs)) / (2.0 * d);
  const S vtot2 = (sR2 + (Rs / nu) * dnu) + Rs * dphidR;
  const S sphi = sR;
  S vmean = sqrt(smax(vtot2 - sphi * sphi, 0.0));
  S vR_s = sR, vz_s = sz, vp_s = sphi;
  if (val(nu) <= 0.0) {
    vmean = Lift<S>::c(0.0);
    vR_s = Lift<S>::c(0.0);
    vz_s = Lift<S>::c(0.0);
    vp_s = Lift<S>::c(0.0);
  }
  const double g1 = normals[3 * i], g2 = normals[3 * i + 1], g3 = normals[3 * i + 2];
  const S vR = g1 * vR_s, vz = g2 * vz_s, vphi = vmean + g3 * vp_s;
  // clockwise cylindrical -> Cartesian (utils.py:144-162)
  const double phi = atan2(ys, xs);
  const double c = cos(phi), s = sin(phi);
  o.wn[0] = Lift<S>::c(Rs * c);
  o.wn[1] = Lift<S>::c(Rs * s);
  o.wn[2] = Lift<S>::c(zs);
  o.wn[3] = vR * c + vphi * s;
  o.wn[4] = vR * s - vphi * c;
  o.wn[5] = vz;
  // rotation curve and orbital time scale at the new ICs (utils.py:182-273)
  const double x0 = Rs * c, y0 = Rs * s;
  const double R = sqrt(x0 * x0 + y0 * y0);
  const double dRt = 1e-3;
  const double Rsafe = fmax(fabs(R), 2.0 * dRt);
  const S dd = (phi_at(p, f, Rsafe + dRt, 0.0, zs) - phi_at(p, f, Rsafe - dRt, 0.0, zs)) / (2.0 * dRt);
  const S vc = sqrt(smax(Rsafe * dd, 0.0));
  const S om = sqrt(smax(fabs(dd) / Rsafe, 1e-20));
  o.T_orb = (2.0 * kPi) / om;
  o.vsc = smin(smax(0.1 * vc, 1.0), 15.0);
}

// One 64-lane wave per star; every lane returns the same values.
template <class S>
__device__ void star_ic(const smx_potential& p, const PotFields<S>& f, const double* w0,
                        const double* normals, int64_t i, int lane, StarIC<S>& o) {
  const double xs = w0[3 * i], ys = w0[3 * i + 1], zs = w0[3 * i + 2];
  const double Rs = sqrt(xs * xs + ys * ys);
  const double zlo = fabs(zs);
  const S I0 = wave_sum(pressure_part_t(p, f, xs, ys, zlo, JEANS_NPTS, lane));
  const S Ip = wave_sum(pressure_part_t(p, f, Rs + JEANS_DR, 0.0, zlo, JEANS_NPTS, lane));
  const S Im = wave_sum(pressure_part_t(p, f, Rs - JEANS_DR, 0.0, zlo, JEANS_NPTS, lane));
  star_ic_finish(p, f, w0, normals, i, I0, Ip, Im, o);
}

// A pressure integral with its derivatives with respect to every potential
// field (PotField order).
struct FieldGrad {
  double v;
  double d[PF_COUNT];
};

__device__ __forceinline__ FieldGrad wave_sum(FieldGrad g) {
  g.v = wave_sum(g.v);
#pragma unroll
  for (int k = 0; k < PF_COUNT; ++k) g.d[k] = wave_sum(g.d[k]);
  return g;
}

// Density and central-difference dPhi/dz of one component at one point,
// with tangents on that component's fields (kept out of line: inlined
// together, the components' dual evaluations exceed the register file).
__device__ __noinline__ void t3_point(const smx_potential& p, double x, double y, double z, double d,
                                      Dual<4>& rho, Dual<4>& dz) {
  const Dual<4> M = dseed<4>(p.t3_M, 0), L = dseed<4>(p.t3_L, 1), a = dseed<4>(p.t3_a, 2),
                b = dseed<4>(p.t3_b, 3);
  const Dual<4> X = dconst<4>(x), Y = dconst<4>(y);
  rho = t3_rho_t(p, M, L, a, b, X, Y, dconst<4>(z));
  dz = ((-p.G) * t3_ph_t(p, M, L, a, b, X, Y, dconst<4>(z + d)) -
        (-p.G) * t3_ph_t(p, M, L, a, b, X, Y, dconst<4>(z - d))) /
       (2.0 * d);
}

__device__ __noinline__ void mn_point(const smx_potential& p, double x, double y, double z, double d,
                                      Dual<3>& rho, Dual<3>& dz) {
  const Dual<3> X = dconst<3>(x), Y = dconst<3>(y);
  rho = mn_rho_t(p, dseed<3>(p.mn_rho0, 0), dseed<3>(p.mn_Rs, 1), dseed<3>(p.mn_Hs, 2), X, Y,
                 dconst<3>(z));
  const Dual<3> m = dseed<3>(p.mn_mGM, 0), R = dseed<3>(p.mn_Rs, 1), H = dseed<3>(p.mn_Hs, 2);
  dz = (mn_phi_t(p, m, R, H, X, Y, dconst<3>(z + d)) - mn_phi_t(p, m, R, H, X, Y, dconst<3>(z - d))) /
       (2.0 * d);
}

__device__ __noinline__ void nfw_v4_point(const smx_potential& p, double x, double y, double z,
                                          double d, Dual<2>& hz, Dual<1>& vrho, Dual<1>& vz) {
  const Dual<2> m = dseed<2>(p.nfw_mGM, 0), R = dseed<2>(p.nfw_Rs, 1);
  hz = (nfw_phi_t(p, m, R, dconst<2>(x), dconst<2>(y), dconst<2>(z + d)) -
        nfw_phi_t(p, m, R, dconst<2>(x), dconst<2>(y), dconst<2>(z - d))) /
       (2.0 * d);
  const Dual<1> M = dseed<1>(p.v4_M, 0), X = dconst<1>(x), Y = dconst<1>(y);
  vrho = v4_rho_t(p, M, X, Y, dconst<1>(z));
  vz = ((-p.G) * v4_ph_t(p, M, X, Y, dconst<1>(z + d)) - (-p.G) * v4_ph_t(p, M, X, Y, dconst<1>(z - d))) /
       (2.0 * d);
}

// this lane's share of pressure_part_t and of all its field derivatives in
// one pass.  Every field belongs to one component (NFW: mGM, Rs; MN: mGM,
// rho0, Rs, Hs; T3: M, L, a, b; V4: M), rho and Phi are sums over
// components, so per point each component's density and potential are
// evaluated once with tangents on its own fields only, and the integrand
// rho dPhi/dz is differentiated by the product rule.
__device__ FieldGrad pressure_part_grad(const smx_potential& p, double x, double y, double zlo,
                                        int npts, int lane) {
  const double hi = 10.0;
  const double step = (hi - zlo) / static_cast<double>(npts - 1);
  const double d = 5e-3;   // dphi_dz_t's step
  FieldGrad acc;
  acc.v = 0.0;
#pragma unroll
  for (int k = 0; k < PF_COUNT; ++k) acc.d[k] = 0.0;
#pragma unroll 1
  for (int i = lane; i < npts; i += 64) {
    const double zi = lin_pt(zlo, hi, step, i, npts);
    Dual<4> tr, tz;
    Dual<3> mr, mz;
    Dual<2> hz;
    Dual<1> vr, vz;
    t3_point(p, x, y, zi, d, tr, tz);
    mn_point(p, x, y, zi, d, mr, mz);
    nfw_v4_point(p, x, y, zi, d, hz, vr, vz);
    const double rho = (mr.v + tr.v) + vr.v;
    const double G = ((hz.v + mz.v) + tz.v) + vz.v;
    double w = 0.0;
    if (i > 0) w += zi - lin_pt(zlo, hi, step, i - 1, npts);
    if (i < npts - 1) w += lin_pt(zlo, hi, step, i + 1, npts) - zi;
    const double c = 0.5 * w;
    acc.v += c * (rho * G);
    acc.d[PF_NFW_MGM] += c * (rho * hz.d[0]);
    acc.d[PF_NFW_RS] += c * (rho * hz.d[1]);
    acc.d[PF_MN_MGM] += c * (rho * mz.d[0]);
    acc.d[PF_MN_RHO0] += c * (mr.d[0] * G);
    acc.d[PF_MN_RS] += c * (mr.d[1] * G + rho * mz.d[1]);
    acc.d[PF_MN_HS] += c * (mr.d[2] * G + rho * mz.d[2]);
#pragma unroll
    for (int k = 0; k < 4; ++k) acc.d[PF_T3_M + k] += c * (tr.d[k] * G + rho * tz.d[k]);
    acc.d[PF_V4_M] += c * (vr.d[0] * G + rho * vz.d[0]);
  }
  return acc;
}

// fields f0 .. f0 + N - 1 of a FieldGrad as a Dual<N> (beyond PF_COUNT: 0)
template <int N>
__device__ __forceinline__ Dual<N> grad_slice(const FieldGrad& g, int f0) {
  Dual<N> r = dconst<N>(g.v);
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = (f0 + k < PF_COUNT) ? g.d[f0 + k] : 0.0;
  return r;
}

// Density target of cell c: mean over the Sobol points of rho(R, z, phi) R,
// times the cell volume (model.py:192-213); this thread's share of the sum.
template <class S>
__device__ __forceinline__ S cell_mass_part(const smx_potential& p, const PotFields<S>& f, double Rc,
                                            double zc, double pc, double dR, double dz, double dphi,
                                            const double* smp, int64_t S_n, int tid, int nthr) {
  S acc = Lift<S>::c(0.0);
  for (int64_t k = tid; k < S_n; k += nthr) {
    const double R = Rc + (smp[3 * k] - 0.5) * dR;
    const double z = zc + (smp[3 * k + 1] - 0.5) * dz;
    const double ph = pc + (smp[3 * k + 2] - 0.5) * dphi;
    acc += rho_total_t(p, f, Lift<S>::c(R * cos(ph)), Lift<S>::c(R * sin(ph)), Lift<S>::c(z)) * R;
  }
  return acc;
}

}  // namespace smx
