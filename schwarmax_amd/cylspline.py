"""CylSpline potential on the GPU (CylindricalSpline.py, CubicSpline.py).

A potential tabulated as azimuthal Fourier modes Phi_m(R, z), m = 0..Mmax, on
a (R, |z|) grid scaled by asinh(./R0), each mode's real and imaginary part a
separable cubic spline with not-a-knot ends.

``get_phi_m`` (CylindricalSpline.py:269-352) builds the tables from a density:
* rho_m(R, z) by an Nphi-point sum over phi (the reference's convention,
  (1/Nphi) sum rho e^{-i m phi});
* its splines, evaluated at the nodes of a Simpson rule on the log-mapped
  (R', z') plane;
* Phi_m(R_i, z_j) = -G sum_nodes w rho_m (Xi_m(+z') + Xi_m(-z')) 2 pi R' J,
  with the Legendre kernel Xi_m = Q_{m-1/2}(chi)/(pi sqrt(R R')), on the GPU
  (``smx_cylspline_phi_m``, one workgroup per (m, R_i, z_j));
* the splines of Phi_m.
The density sum, spline solves and node values are small host steps that
run once per potential.

``evaluate_phi``, ``evaluate_phi_axisymmetric``, ``get_acc`` and
``get_density`` (CylindricalSpline.py:355-536) evaluate the tables at points
on the GPU (``smx_cylspline_eval``): Phi, -grad Phi from the spline's analytic
derivatives, and nabla^2 Phi/(4 pi G) from its analytic second derivatives
(the trace the reference takes with jax.hessian).

Q_{m-1/2} is computed to ~1e-14 from its hypergeometric series (the
reference uses Pade and truncated-series approximations with the same
1e-12 floor on the logarithm's argument), so tables agree with the
reference's to that approximation's accuracy.
"""
import math

import numpy as np
import torch

from . import _lib
from .potentials import G

MAX_MODES = _lib.CYLSPLINE_MAX_MODES


def simpson_weights(n):
    """Composite Simpson weights on [0, 1], n odd (CylindricalSpline.py:193-199)."""
    w = np.ones(n)
    w[1:-1:2] = 4.0
    w[2:-1:2] = 2.0
    return w * ((1.0 / (n - 1)) / 3.0)


def _not_a_knot_matrix(x):
    """Second-derivative system of the not-a-knot cubic spline on knots x
    (CubicSpline.py:5-53): third derivative continuous at x_1 and x_{n-2}."""
    n = x.shape[0]
    if n < 4:
        raise ValueError("a not-a-knot spline needs at least 4 knots")
    h = np.diff(x)
    A = np.zeros((n, n))
    A[0, :3] = [h[1], -(h[0] + h[1]), h[0]]
    for i in range(1, n - 1):
        A[i, i - 1:i + 2] = [h[i - 1], 2.0 * (h[i - 1] + h[i]), h[i]]
    A[n - 1, n - 3:] = [h[n - 2], -(h[n - 3] + h[n - 2]), h[n - 3]]
    return A, h


def spline_second_derivatives(x, y):
    """M with A M = b for each column of y (n, ...) along axis 0."""
    A, h = _not_a_knot_matrix(np.asarray(x, dtype=np.float64))
    y = np.asarray(y, dtype=np.float64)
    flat = y.reshape(y.shape[0], -1)
    b = np.zeros_like(flat)
    b[1:-1] = 6.0 * ((flat[2:] - flat[1:-1]) / h[1:, None] - (flat[1:-1] - flat[:-2]) / h[:-1, None])
    return np.linalg.solve(A, b).reshape(y.shape)


def precompute_splines(grid, values):
    """(M_x, M_y): second derivatives along x for each y, along y for each x
    (CubicSpline.py:55-99)."""
    gx, gy = grid
    v = np.asarray(values, dtype=np.float64)
    return spline_second_derivatives(gx, v), spline_second_derivatives(gy, v.T).T


def spline_eval(u, v, grid, values, M_x, M_y, fill_value=0.0):
    """Separable cubic interpolant (CubicSpline.py:102-199) at points (u, v)."""
    gx, gy = (np.asarray(g, dtype=np.float64) for g in grid)
    u = np.asarray(u, dtype=np.float64)
    v = np.asarray(v, dtype=np.float64)
    oob = (u < gx[0]) | (u > gx[-1]) | (v < gy[0]) | (v > gy[-1])
    uc, vc = np.clip(u, gx[0], gx[-1]), np.clip(v, gy[0], gy[-1])
    i = np.clip(np.searchsorted(gx, uc) - 1, 0, gx.size - 2)
    j = np.clip(np.searchsorted(gy, vc) - 1, 0, gy.size - 2)
    hx, hy = gx[i + 1] - gx[i], gy[j + 1] - gy[j]
    tx, ty = (uc - gx[i]) / hx, (vc - gy[j]) / hy
    sx, sy = 1.0 - tx, 1.0 - ty

    def along_x(z0, z1, m0, m1):
        return sx * z0 + tx * z1 + ((sx ** 3 - sx) * m0 + (tx ** 3 - tx) * m1) * hx ** 2 / 6.0
    f0 = along_x(values[i, j], values[i + 1, j], M_x[i, j], M_x[i + 1, j])
    f1 = along_x(values[i, j + 1], values[i + 1, j + 1], M_x[i, j + 1], M_x[i + 1, j + 1])
    my0 = sx * M_y[i, j] + tx * M_y[i + 1, j]
    my1 = sx * M_y[i, j + 1] + tx * M_y[i + 1, j + 1]
    out = sy * f0 + ty * f1 + ((sy ** 3 - sy) * my0 + (ty ** 3 - ty) * my1) * hy ** 2 / 6.0
    return np.where(oob, fill_value, out)


def _host(a):
    if isinstance(a, torch.Tensor):
        return a.detach().to('cpu', torch.float64).numpy()
    return np.asarray(a, dtype=np.float64)


def quadrature_nodes(R1, Rmax, z1, Zmax, N_int):
    """Simpson nodes of the log-mapped (R', z') plane: R' = R1((1 + Rmax/R1)^xi - 1),
    z' likewise, xi and eta on odd grids of max(9, floor(sqrt(max(16, N_int))))
    points (rounded up to odd); returns R', z' and weight * 2 pi R' * Jacobian,
    flattened xi-major (CylindricalSpline.py:201-222, 312-324)."""
    base = max(9, int(np.sqrt(max(16, N_int))))
    base += abs(base % 2 - 1)
    w = simpson_weights(base)
    t = np.linspace(0.0, 1.0, base)
    XI, ETA = np.meshgrid(t, t, indexing='ij')
    LR, LZ = math.log(1.0 + Rmax / R1), math.log(1.0 + Zmax / z1)
    Rp = R1 * (np.power(1.0 + Rmax / R1, XI) - 1.0)
    zp = z1 * (np.power(1.0 + Zmax / z1, ETA) - 1.0)
    J = (LR * (R1 + Rp)) * (LZ * (z1 + zp))
    wk = np.outer(w, w) * (2.0 * math.pi) * Rp * J
    return Rp.ravel(), zp.ravel(), wk.ravel(), base


def get_phi_m(rho_fn, params, NR, NZ, Rmin, Rmax, Zmin, Zmax, Mmax, Nphi, N_int):
    """CylindricalSpline.py:269-352: mode tables of the potential of rho_fn.

    rho_fn(x, y, z, params) is called once on (NR, NZ+1, Nphi) NumPy arrays
    (z >= 0, the reference's even-in-z assumption) and may return NumPy or
    torch.  Returns the reference's dict (NumPy arrays): Rgrid, Zgrid (scaled
    grids), R0, m, Phi_m_real, Phi_m_img, Mx_real, My_real, Mx_img, My_img.
    """
    Mmax, NR, NZ, Nphi = int(Mmax), int(NR), int(NZ), int(Nphi)
    if not 0 <= Mmax < MAX_MODES:
        raise ValueError("Mmax must lie in [0, %d]" % (MAX_MODES - 1))
    if NR < 4 or NZ + 1 < 4:
        raise ValueError("the splines need NR >= 4 and NZ >= 3")
    M = np.arange(Mmax + 1)
    R = np.geomspace(max(Rmin, 1e-3), Rmax, NR)
    Zn = np.concatenate([[0.0], np.geomspace(max(Zmin, 1e-3), Zmax, NZ)])
    R0 = R[NR // 2]
    phi = np.linspace(0.0, 2 * np.pi, Nphi, endpoint=False)
    Rg, Zg = np.meshgrid(R, Zn, indexing='ij')
    x = Rg[..., None] * np.cos(phi)
    y = Rg[..., None] * np.sin(phi)
    z = np.broadcast_to(np.abs(Zg)[..., None], x.shape).copy()
    rho = _host(rho_fn(x, y, z, params)).reshape(x.shape)
    E = np.exp(-1j * M[:, None] * phi[None, :])
    rho_m = np.einsum('ijk,mk->mij', rho, E) * ((2.0 * np.pi / Nphi) / (2.0 * np.pi))
    u, v = np.arcsinh(R / R0), np.arcsinh(Zn / R0)

    Rp, zp, wk, _ = quadrature_nodes(R[1], Rmax, Zn[1], Zmax, N_int)
    un, vn = np.arcsinh(Rp / R0), np.arcsinh(np.abs(zp) / R0)
    nodes = np.empty((Mmax + 1, 2, Rp.size))
    for m in M:
        for part, vals in enumerate((rho_m[m].real, rho_m[m].imag)):
            mx, my = precompute_splines((u, v), vals)
            nodes[m, part] = spline_eval(un, vn, (u, v), vals, mx, my, fill_value=0.0)

    L = _lib.lib()
    dev = _lib.dev
    out = torch.empty((Mmax + 1, 2, NR, NZ + 1), dtype=torch.float64, device='cuda')
    Rt, Zt, nR_, nz_, nw_, rn_ = (dev(a) for a in (R, Zn, Rp, zp, wk, nodes))
    _lib.check(L.smx_cylspline_phi_m(Mmax + 1, _lib.ptr(Rt), NR, _lib.ptr(Zt), NZ + 1,
                                     _lib.ptr(nR_), _lib.ptr(nz_), _lib.ptr(nw_), Rp.size,
                                     _lib.ptr(rn_), float(G), _lib.ptr(out), _lib.stream_ptr()),
               'smx_cylspline_phi_m')
    phi_m = out.cpu().numpy()
    tab = {k: np.zeros((Mmax + 1, NR, NZ + 1)) for k in
           ('Phi_m_real', 'Phi_m_img', 'Mx_real', 'My_real', 'Mx_img', 'My_img')}
    for m in M:
        for part, (kv, kx, ky) in enumerate((('Phi_m_real', 'Mx_real', 'My_real'),
                                            ('Phi_m_img', 'Mx_img', 'My_img'))):
            tab[kv][m] = phi_m[m, part]
            tab[kx][m], tab[ky][m] = precompute_splines((u, v), phi_m[m, part])
    return {'Rgrid': u, 'Zgrid': v, 'R0': np.array([R0]), 'm': M, **tab}


class _Table:
    """The mode tables of one dict in the device layout of smx_cylspline."""

    def __init__(self, d):
        self.u = _lib.dev(d['Rgrid'])
        self.v = _lib.dev(d['Zgrid'])
        parts = [np.asarray(d[k], dtype=np.float64) for k in
                 ('Phi_m_real', 'Phi_m_img', 'Mx_real', 'Mx_img', 'My_real', 'My_img')]
        coef = np.stack([p.transpose(1, 2, 0) for p in parts], axis=-1)   # (nR, nZ, NM, 6)
        self.coef = _lib.dev(coef)
        self.nR, self.nZ, self.n_modes = coef.shape[0], coef.shape[1], coef.shape[2]
        self.R0 = float(np.asarray(d['R0']).ravel()[0])
        if self.n_modes > MAX_MODES:
            raise ValueError("at most %d modes" % MAX_MODES)
        if self.u.numel() != self.nR or self.v.numel() != self.nZ or self.nR < 2 or self.nZ < 2:
            raise ValueError("Rgrid/Zgrid lengths must match the mode tables (at least 2 each)")

    def struct(self, n_eval):
        return _lib.SmxCylspline(self.u.data_ptr(), self.v.data_ptr(), self.coef.data_ptr(),
                                 self.nR, self.nZ, self.n_modes, int(n_eval), self.R0, float(G))


def _table(d):
    t = d.get('_smx_table')
    if t is None:
        t = _Table(d)
        d['_smx_table'] = t
    return t


def evaluate(x, y, z, dict_phi_m, modes=None, phi=True, acc=False, density=False):
    """(phi, acc (..., 3), rho) at points from one device launch; modes limits
    the sum to m < modes."""
    from .potentials import _points
    t = _table(dict_phi_m)
    xyz, shape = _points(x, y, z)
    n = xyz.numel() // 3
    outs = [torch.empty(shape, dtype=torch.float64, device='cuda') if phi else None,
            torch.empty(shape + (3,), dtype=torch.float64, device='cuda') if acc else None,
            torch.empty(shape, dtype=torch.float64, device='cuda') if density else None]
    st = t.struct(t.n_modes if modes is None else modes)
    _lib.check(_lib.lib().smx_cylspline_eval(st, _lib.ptr(xyz), n, *[_lib.ptr(o) for o in outs],
                                             _lib.stream_ptr()), 'smx_cylspline_eval')
    return outs


def evaluate_phi(x, y, z, dict_phi_m):
    """CylindricalSpline.py:355-386: sum over modes of f_m Re(Phi_m e^{i m phi})."""
    return evaluate(x, y, z, dict_phi_m)[0]


def evaluate_phi_axisymmetric(x, y, z, dict_phi_m):
    """CylindricalSpline.py:389-415: the m = 0 mode."""
    return evaluate(x, y, z, dict_phi_m, modes=1)[0]


def get_acc(x, y, z, params, eps=5e-4):
    """CylindricalSpline.py:489-520: -grad Phi from the spline's analytic
    derivatives, shaped (3, ...) like the reference's output (eps unused
    there too)."""
    a = evaluate(x, y, z, params, phi=False, acc=True)[1]
    return a.movedim(-1, 0)


def get_density(x, y, z, params):
    """CylindricalSpline.py:529-536: nabla^2 Phi / (4 pi G)."""
    return evaluate(x, y, z, params, phi=False, density=True)[2]
