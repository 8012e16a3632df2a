"""CPU FP64 restatement of the CylSpline potential (CylindricalSpline.py,
CubicSpline.py).  Test infrastructure only: the checker for
schwarmax_amd.cylspline, never called by it.

Follows the reference's formulas term by term:
* Q_{m-1/2}(chi) = sqrt(pi) Gamma(m+1/2) / (Gamma(m+1) 2^{m+1/2}) chi^{-m-1/2}
  2F1((2m+1)/4, (2m+3)/4; m+1; 1/chi^2) (jax_legendreQ, CylindricalSpline.py:
  145-160), with scipy's hyp2f1 in place of the reference's Pade tables for
  1 - 1/chi^2 >= 0.05 and, nearer chi = 1, the logarithmic series the
  reference truncates at (1 - 1/chi^2)^4, summed to convergence with the
  logarithm's argument floored at 1e-12 as the reference floors it;
* the not-a-knot spline solved column by column as a dense system
  (CubicSpline.py:5-99) and the separable interpolant with derivatives
  (CubicSpline.py:202-305);
* get_phi_m's grids, density modes, Simpson nodes and quadrature
  (CylindricalSpline.py:193-352), evaluate_phi and get_acc
  (CylindricalSpline.py:355-520);
* get_density as -div(acc)/(4 pi G) by central differences of the analytic
  acceleration (the reference traces jax.hessian).
"""
import math

import numpy as np
from scipy.special import digamma, gamma, hyp2f1

from .constants import G


def _hyp_near_one(a, b, y):
    """F(a, b; a+b; 1-y) by Abramowitz & Stegun 15.3.10, summed to convergence,
    with the logarithm's argument floored at 1e-12 as the reference's
    near-one branch does (CylindricalSpline.py:112, 129-133)."""
    L = np.log(np.maximum(y, 1e-12))
    pa, pb, p1 = digamma(a), digamma(b), digamma(1.0)
    c, yn, s = 1.0, np.ones_like(y), np.zeros_like(y)
    for n in range(60):
        s = s + c * (2.0 * p1 - pa - pb - L) * yn
        c *= (a + n) * (b + n) / ((n + 1.0) * (n + 1.0))
        p1 += 1.0 / (n + 1.0)
        pa += 1.0 / (a + n)
        pb += 1.0 / (b + n)
        yn = yn * y
    return gamma(a + b) / (gamma(a) * gamma(b)) * s


def legendre_q_half(m, chi):
    chi = np.maximum(np.asarray(chi, dtype=np.float64), 1.0)
    x = 1.0 / (chi * chi)
    y = 1.0 - x
    a, b = (2 * m + 1) / 4.0, (2 * m + 3) / 4.0
    pref = math.sqrt(math.pi) * gamma(m + 0.5) / (gamma(m + 1.0) * 2.0 ** (m + 0.5))
    with np.errstate(divide='ignore'):
        near = _hyp_near_one(a, b, np.minimum(y, 0.05))
    F = np.where(y < 0.05, near, hyp2f1(a, b, m + 1.0, np.minimum(x, 0.95)))
    return pref * chi ** (-m - 0.5) * F


def kernel_xi(m, R, z, Rp, zp):
    """jax_kernel_Xi_m (CylindricalSpline.py:162-190)."""
    R, z, Rp, zp = np.broadcast_arrays(*(np.asarray(a, dtype=np.float64) for a in (R, z, Rp, zp)))
    dz = z - zp
    val_zero = (1.0 / np.sqrt(R * R + Rp * Rp + dz * dz)) if m == 0 else np.zeros(R.shape)
    with np.errstate(divide='ignore', invalid='ignore'):
        chi = np.maximum((R * R + Rp * Rp + dz * dz) / (2.0 * R * Rp), 1.0)
        val = legendre_q_half(m, chi) / (np.pi * np.sqrt(R * Rp))
    return np.where((R < 1e-3) | (Rp < 1e-3), val_zero, val)


def second_derivatives(x, y):
    """jax_compute_second_derivatives (CubicSpline.py:5-53)."""
    n = len(x)
    h = np.diff(x)
    A = np.zeros((n, n))
    b = np.zeros(n)
    A[0, 0], A[0, 1], A[0, 2] = h[1], -(h[0] + h[1]), h[0]
    for i in range(1, n - 1):
        A[i, i - 1], A[i, i], A[i, i + 1] = h[i - 1], 2.0 * (h[i - 1] + h[i]), h[i]
        b[i] = 6.0 * ((y[i + 1] - y[i]) / h[i] - (y[i] - y[i - 1]) / h[i - 1])
    A[n - 1, n - 3], A[n - 1, n - 2], A[n - 1, n - 1] = h[n - 2], -(h[n - 3] + h[n - 2]), h[n - 3]
    return np.linalg.solve(A, b)


def precompute_splines(grid, values):
    """jax_precompute_splines (CubicSpline.py:55-99)."""
    nx, ny = values.shape
    M_x = np.zeros((nx, ny))
    M_y = np.zeros((nx, ny))
    for j in range(ny):
        M_x[:, j] = second_derivatives(grid[0], values[:, j])
    for i in range(nx):
        M_y[i, :] = second_derivatives(grid[1], values[i, :])
    return M_x, M_y


def spline_eval_with_derivs(xi, grid, values, M_x, M_y, fill_value=np.nan):
    """cubic_spline_evaluate_with_derivs (CubicSpline.py:202-305): value and
    its derivatives along the two grid coordinates."""
    xi = np.asarray(xi, dtype=np.float64).reshape(-1, 2)
    xp, yp = xi[:, 0], xi[:, 1]
    gx, gy = grid
    oob = (xp < gx[0]) | (xp > gx[-1]) | (yp < gy[0]) | (yp > gy[-1])
    xp = np.clip(xp, gx[0], gx[-1])
    yp = np.clip(yp, gy[0], gy[-1])
    ix = np.clip(np.searchsorted(gx, xp) - 1, 0, len(gx) - 2)
    iy = np.clip(np.searchsorted(gy, yp) - 1, 0, len(gy) - 2)
    hx = gx[ix + 1] - gx[ix]
    hy = gy[iy + 1] - gy[iy]
    tx = (xp - gx[ix]) / hx
    ty = (yp - gy[iy]) / hy
    omx, omy = 1.0 - tx, 1.0 - ty
    z00, z10, z01, z11 = (values[ix, iy], values[ix + 1, iy], values[ix, iy + 1],
                          values[ix + 1, iy + 1])
    a00, a10, a01, a11 = M_x[ix, iy], M_x[ix + 1, iy], M_x[ix, iy + 1], M_x[ix + 1, iy + 1]
    f0 = omx * z00 + tx * z10 + ((omx ** 3 - omx) * a00 + (tx ** 3 - tx) * a10) * hx ** 2 / 6.0
    f1 = omx * z01 + tx * z11 + ((omx ** 3 - omx) * a01 + (tx ** 3 - tx) * a11) * hx ** 2 / 6.0
    d0 = (z10 - z00) + ((-3.0 * omx ** 2 + 1.0) * a00 + (3.0 * tx ** 2 - 1.0) * a10) * hx ** 2 / 6.0
    d1 = (z11 - z01) + ((-3.0 * omx ** 2 + 1.0) * a01 + (3.0 * tx ** 2 - 1.0) * a11) * hx ** 2 / 6.0
    b00, b10, b01, b11 = M_y[ix, iy], M_y[ix + 1, iy], M_y[ix, iy + 1], M_y[ix + 1, iy + 1]
    m0 = omx * b00 + tx * b10
    m1 = omx * b01 + tx * b11
    val = omy * f0 + ty * f1 + ((omy ** 3 - omy) * m0 + (ty ** 3 - ty) * m1) * hy ** 2 / 6.0
    dtx = (omy * d0 + ty * d1 + ((omy ** 3 - omy) * (b10 - b00) + (ty ** 3 - ty) * (b11 - b01))
           * hy ** 2 / 6.0)
    dty = (f1 - f0) + ((-3.0 * omy ** 2 + 1.0) * m0 + (3.0 * ty ** 2 - 1.0) * m1) * hy ** 2 / 6.0
    return (np.where(oob, fill_value, val), np.where(oob, 0.0, dtx / hx),
            np.where(oob, 0.0, dty / hy))


def simpson_weights(n):
    w = np.ones(n)
    w[1:-1:2] = 4.0
    w[2:-1:2] = 2.0
    return w * (1.0 / (n - 1)) / 3.0


def get_phi_m(rho_fn, params, NR, NZ, Rmin, Rmax, Zmin, Zmax, Mmax, Nphi, N_int):
    """CylindricalSpline.py:269-352."""
    M = np.arange(Mmax + 1)
    R = np.geomspace(max(Rmin, 1e-3), Rmax, NR)
    Z = np.concatenate([[0.0], np.geomspace(max(Zmin, 1e-3), Zmax, NZ)])
    R0 = R[NR // 2]
    Rg, Zg = np.meshgrid(R, Z, indexing='ij')
    phi = np.linspace(0.0, 2 * np.pi, Nphi, endpoint=False)
    zz = np.broadcast_to(np.abs(Zg)[..., None], Rg.shape + (Nphi,))
    rho_phi = rho_fn(Rg[..., None] * np.cos(phi), Rg[..., None] * np.sin(phi), zz, params)
    e = np.exp(-1j * M[:, None] * phi[None, :])
    rho_m = np.sum(rho_phi[..., None, :] * e[None, None], axis=-1) * (2.0 * np.pi / Nphi) / (2.0 * np.pi)
    rho_m = rho_m.transpose(2, 0, 1)
    gR, gZ = np.arcsinh(R / R0), np.arcsinh(Z / R0)

    base = max(9, int(np.sqrt(max(16, N_int))))
    base += abs(base % 2 - 1)
    W2D = np.einsum('i,j->ij', simpson_weights(base), simpson_weights(base))
    t = np.linspace(0.0, 1.0, base)
    XI, ETA = np.meshgrid(t, t, indexing='ij')
    Rmn, Rmx, zmn, zmx = R[1], Rmax, Z[1], Zmax
    LR, LZ = np.log(1.0 + Rmx / Rmn), np.log(1.0 + zmx / zmn)
    Rp = Rmn * (np.power(1.0 + Rmx / Rmn, XI) - 1.0)
    zp = zmn * (np.power(1.0 + zmx / zmn, ETA) - 1.0)
    Jmap = (LR * (Rmn + Rp)) * (LZ * (zmn + zp))
    pts = np.column_stack([np.arcsinh(Rp / R0).ravel(), np.arcsinh(np.abs(zp) / R0).ravel()])

    out = {k: np.zeros((Mmax + 1, NR, NZ + 1)) for k in
           ('Phi_m_real', 'Phi_m_img', 'Mx_real', 'My_real', 'Mx_img', 'My_img')}
    Rt, Zt = Rg.ravel()[:, None], Zg.ravel()[:, None]
    for m in M:
        rg = 0j
        for part, vals in ((1.0, rho_m[m].real), (1j, rho_m[m].imag)):
            mx, my = precompute_splines((gR, gZ), vals)
            rg = rg + part * spline_eval_with_derivs(pts, (gR, gZ), vals, mx, my, 0.0)[0]
        Xi = (kernel_xi(m, Rt, Zt, Rp.ravel()[None], zp.ravel()[None])
              + kernel_xi(m, Rt, Zt, Rp.ravel()[None], -zp.ravel()[None]))
        F = rg[None] * Xi * (2.0 * np.pi) * Rp.ravel()[None] * Jmap.ravel()[None]
        ph = (-G * np.sum(W2D.ravel()[None] * F, axis=1)).reshape(NR, NZ + 1)
        for vals, kv, kx, ky in ((ph.real, 'Phi_m_real', 'Mx_real', 'My_real'),
                                 (ph.imag, 'Phi_m_img', 'Mx_img', 'My_img')):
            out[kv][m] = vals
            out[kx][m], out[ky][m] = precompute_splines((gR, gZ), vals)
    return {'Rgrid': gR, 'Zgrid': gZ, 'R0': np.array([R0]), 'm': M, **out}


def _phi_and_scaled_derivs(x, y, z, d, modes=None):
    """_evaluate_phi_and_scaled_derivs (CylindricalSpline.py:418-479)."""
    x, y, z = (np.asarray(a, dtype=np.float64).ravel() for a in (x, y, z))
    R = np.sqrt(x * x + y * y)
    ph = np.arctan2(y, x)
    za = np.abs(z)
    R0 = d['R0'][0]
    pts = np.column_stack([np.arcsinh(R / R0), np.arcsinh(za / R0)])
    grid = (d['Rgrid'], d['Zgrid'])
    acc = [np.zeros_like(R) for _ in range(4)]
    for m in (d['m'] if modes is None else d['m'][:modes]):
        rv, rdx, rdy = spline_eval_with_derivs(pts, grid, d['Phi_m_real'][m], d['Mx_real'][m],
                                               d['My_real'][m], 0.0)
        iv, idx_, idy = spline_eval_with_derivs(pts, grid, d['Phi_m_img'][m], d['Mx_img'][m],
                                                d['My_img'][m], 0.0)
        c, s = np.cos(m * ph), np.sin(m * ph)
        f = 1.0 if m == 0 else 2.0
        acc[0] += f * (rv * c - iv * s)
        acc[1] += f * (rdx * c - idx_ * s)
        acc[2] += f * (rdy * c - idy * s)
        acc[3] += f * (-m * rv * s - m * iv * c)
    return acc[0], acc[1], acc[2], acc[3], R, R0, za


def evaluate_phi(x, y, z, d, modes=None):
    """CylindricalSpline.py:355-386 (modes=1: evaluate_phi_axisymmetric)."""
    return _phi_and_scaled_derivs(x, y, z, d, modes)[0]


def get_acc(x, y, z, d):
    """CylindricalSpline.py:489-520; returns (3, n)."""
    z = np.asarray(z, dtype=np.float64).ravel()
    x = np.asarray(x, dtype=np.float64).ravel()
    y = np.asarray(y, dtype=np.float64).ravel()
    _, dRs, dzs, dph, R, R0, za = _phi_and_scaled_derivs(x, y, z, d)
    dphi_dR = dRs / np.sqrt(R * R + R0 * R0)
    dphi_dz = dzs / np.sqrt(za * za + R0 * R0) * np.sign(z)
    tiny = 1e-12
    with np.errstate(divide='ignore', invalid='ignore'):
        invR = np.where(R > tiny, 1.0 / R, 0.0)
    dRdx = np.where(R > tiny, x * invR, 0.0)
    dRdy = np.where(R > tiny, y * invR, 0.0)
    dpdx = np.where(R > tiny, -y * invR * invR, 0.0)
    dpdy = np.where(R > tiny, x * invR * invR, 0.0)
    return -np.array([dphi_dR * dRdx + dph * dpdx, dphi_dR * dRdy + dph * dpdy, dphi_dz])


def get_density(x, y, z, d, h=1e-5):
    """nabla^2 Phi / (4 pi G) as -div(acc) by central differences."""
    x, y, z = (np.asarray(a, dtype=np.float64).ravel() for a in (x, y, z))
    div = np.zeros_like(x)
    for k in range(3):
        e = [np.zeros_like(x) for _ in range(3)]
        e[k] = np.full_like(x, h)
        ap = get_acc(x + e[0], y + e[1], z + e[2], d)[k]
        am = get_acc(x - e[0], y - e[1], z - e[2], d)[k]
        div += (ap - am) / (2.0 * h)
    return -div / (4.0 * np.pi * G)
