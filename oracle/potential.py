"""Canonical SchwarMAX potential and density, FP64, elementwise over arrays.

Restates potentials.py:8-65 (NFW, Miyamoto-Nagai), dehnen_bar.py:104-157,
211-241, 269-278, 913-933, 1254-1260, 1318-1326, 1371-1383, 1425-1438,
1526-1543 (Dehnen & Aly T3 bar and V4 bulge as needle convolutions) and
likelihood.py:10-71 (the canonical ``density_func`` / ``potential_func``).
Every function accepts NumPy arrays or :class:`oracle.dual.Dual` values, so the
same code yields Phi and, through dual numbers, the exact gradient that
``jax.grad`` gives the reference (model.py:152-157).
"""
import numpy as np

from .constants import G, EPSILON
from . import dual as D

V4_A, V4_B, V4_L, V4_GAMMA = 0.5, 0.5, 0.1, 0.0      # likelihood.py:10
GAMMA_BAR = 1.0                                      # likelihood.py:11


def get_mat(dirx, diry, dirz):
    """Rodrigues rotation taking e_z onto dir (utils.py:70-99)."""
    v1 = np.array([0.0, 0.0, 1.0])
    v2 = np.array([dirx, diry, dirz], dtype=np.float64)
    v2 = v2 / (np.linalg.norm(v2) + EPSILON)
    angle = np.arccos(np.clip(np.dot(v1, v2), -1.0, 1.0))
    v3 = np.cross(v1, v2)
    v3 = v3 / (np.linalg.norm(v3) + EPSILON)
    K = np.array([[0, -v3[2], v3[1]], [v3[2], 0, -v3[0]], [-v3[1], v3[0], 0]])
    return np.eye(3) + np.sin(angle) * K + (1 - np.cos(angle)) * (K @ K)


def _shift_rotate(x, y, z, p):
    """_shift then _rotate (potentials.py:8-15)."""
    R = get_mat(p['dirx'], p['diry'], p['dirz'])
    dx, dy, dz = x - p['x_origin'], y - p['y_origin'], z - p['z_origin']
    return [R[i, 0] * dx + R[i, 1] * dy + R[i, 2] * dz for i in range(3)]


def nfw_potential(x, y, z, p):
    """potentials.py:18-27."""
    rx, ry, rz = _shift_rotate(x, y, z, p)
    r = D.sqrt((rx / p['a']) ** 2 + (ry / p['b']) ** 2 + (rz / p['c']) ** 2 + EPSILON)
    return -G * 10.0 ** p['logM'] * D.log(1 + r / p['Rs']) / (r + EPSILON)


def mn_density(x, y, z, p):
    """potentials.py:29-49."""
    rx, ry, rz = [c + EPSILON for c in _shift_rotate(x, y, z, p)]
    R = D.sqrt(rx ** 2 + ry ** 2)
    beta = D.sqrt(rz ** 2 + p['Hs_disc'] ** 2)
    D2 = R * R + (p['Rs_disc'] + beta) ** 2
    num = p['Rs_disc'] * R * R + (p['Rs_disc'] + 3.0 * beta) * (p['Rs_disc'] + beta) ** 2
    den = beta ** 3 * D2 ** 2.5
    return (p['Hs_disc'] ** 2 * 10.0 ** p['logM_disc'] / (4 * np.pi)) * (num / den)


def mn_potential(x, y, z, p):
    """potentials.py:51-65."""
    rx, ry, rz = [c + EPSILON for c in _shift_rotate(x, y, z, p)]
    R = (rx ** 2 + ry ** 2) ** 0.5
    denom2 = R ** 2 + (p['Rs_disc'] + (rz ** 2 + p['Hs_disc'] ** 2) ** 0.5) ** 2
    return -G * 10.0 ** p['logM_disc'] / (denom2 ** 0.5)


# --------------------------------------------------------------------------
# Dehnen & Aly needle-convolved kernels
# --------------------------------------------------------------------------
def compute_ACRn(x, u2, R2, n):
    """A_n, B_n, Ax_n by the two-step recurrence (dehnen_bar.py:104-157)."""
    iR2 = 1.0 / R2
    iU2 = 1.0 / u2
    r = D.sqrt(R2)
    if n % 2 == 1:
        Ax = 1.0 / r
        A = D.log(x + r)
        B = r
        cur = 1
        if n >= 3:
            A, B, Ax = ((cur - 1) * A + x * Ax) / (cur * u2), -Ax / cur, Ax * iR2
            cur = 3
        while cur < n:
            A, B, Ax = ((cur - 1) * A + x * Ax) / (cur * u2), -Ax / cur, Ax * iR2
            cur += 2
    else:
        iU = D.sqrt(iU2)
        Ax = iR2
        A = iU * D.arctan(iU * x)
        B = 0.5 * D.log(R2)
        cur = 2
        while cur < n:
            A, B, Ax = ((cur - 1) * A + x * Ax) / (cur * u2), -Ax / cur, Ax * iR2
            cur += 2
    return A, B, Ax


def acrn_A_C(x, y, z, n):
    """(A_n, C_n = x A_n - B_n) (dehnen_bar.py:269-278)."""
    u2 = y * y + z * z
    R2 = x * x + u2
    A, B, _ = compute_ACRn(x, u2, R2, n)
    return A, x * A - B


def convolve_val(M, x, L, gamma, y, z, n):
    """Needle convolution (dehnen_bar.py:211-241)."""
    c1 = M * (1 - gamma) / (2 * L)
    c2 = M * gamma / (L * L)
    Ap, Cp = acrn_A_C(x + L, y, z, n)
    Am, Cm = acrn_A_C(x - L, y, z, n)
    result = c1 * (Ap - Am)
    result = result + c2 * (Cp + Cm - 2 * acrn_A_C(x, y, z, n)[1])
    return result


def rfunc_axi(M, x, y, z, n):
    """M / r^n (dehnen_bar.py:913-925)."""
    iR2 = 1.0 / (x * x + y * y + z * z)
    if n % 2 == 1:
        F = M * D.sqrt(iR2)
        cur = 1
    else:
        F = M * iR2
        cur = 2
    while cur < n:
        F = F * iR2
        cur += 2
    return F


def rn_cond(M, x, y, z, n, L, gamma):
    """Barred / axisymmetric dispatch on L > 0 (dehnen_bar.py:1254-1260)."""
    if L > 0:
        return convolve_val(M, x, L, gamma, y, z, n)     # _rfunc_bar, 928-933
    return rfunc_axi(M, x, y, z, n)


def t3_potential(x, y, z, M, a, b, L, gamma):
    """dehnen_bar.py:1318-1326."""
    ze = D.sqrt(z * z + b * b)
    Z = ze + a
    aZ = a * Z
    ph = rn_cond(M, x, y, Z, 1, L, gamma)
    ph = ph + a * (Z - a / 3) * rn_cond(M, x, y, Z, 3, L, gamma)
    ph = ph + aZ ** 2 * rn_cond(M, x, y, Z, 5, L, gamma)
    return -G * ph


def v4_potential(x, y, z, M, a, b, L, gamma):
    """dehnen_bar.py:1371-1383."""
    ze = D.sqrt(z * z + b * b)
    Z = ze + a
    aZ = a * Z
    hB2 = 0.5 * b * b
    ph = rn_cond(M, x, y, Z, 1, L, gamma)
    ph = ph + (a * (Z - 0.4 * a) + hB2) * rn_cond(M, x, y, Z, 3, L, gamma)
    ph = ph + 3 * a * (0.2 * aZ * (Z + Z - a) + hB2 * (Z - 0.2 * a)) * rn_cond(M, x, y, Z, 5, L, gamma)
    ph = ph + aZ ** 2 * (aZ + 6 * hB2) * rn_cond(M, x, y, Z, 7, L, gamma)
    ph = ph + (7 * hB2 * aZ ** 3 * Z / ze) * rn_cond(M, x, y, Z, 9, L, gamma)
    return -G * ph


def t1_potential(x, y, z, M, a, b, L, gamma):
    """dehnen_bar.py:1303-1307 (used only as a test pin: T1 = Miyamoto-Nagai)."""
    ze = D.sqrt(z * z + b * b)
    return -G * rn_cond(M, x, y, ze + a, 1, L, gamma)


def t3_density(x, y, z, M, a, b, L, gamma):
    """dehnen_bar.py:1425-1438."""
    ze = D.sqrt(z * z + b * b)
    iz = 1.0 / ze
    Z = ze + a
    Z2 = Z * Z
    A2 = a * a
    rh = 3 * ze ** 3 * rn_cond(M, x, y, Z, 5, L, gamma)
    rh = rh + 5 * a * Z2 * ((3 * ze - 2 * a) * ze + A2) * rn_cond(M, x, y, Z, 7, L, gamma)
    rh = rh + 35 * A2 * Z2 * Z2 * ze * rn_cond(M, x, y, Z, 9, L, gamma)
    return (b * b * iz ** 3) * (0.25 / np.pi) * rh


def v4_density(x, y, z, M, a, b, L, gamma):
    """dehnen_bar.py:1526-1543."""
    ze = D.sqrt(z * z + b * b)
    iz = 1.0 / ze
    Z = ze + a
    Z2 = Z * Z
    A2 = a * a
    AZ = a * Z
    rh = 5 * ze ** 3 * rn_cond(M, x, y, Z, 7, L, gamma)
    rh = rh + 7 * a * Z2 * (5 * ze * ze - 4 * a * ze + 3 * A2 - 2 * a * A2 * iz
                            + A2 * A2 * iz * iz) * rn_cond(M, x, y, Z, 9, L, gamma)
    rh = rh + 63 * A2 * Z2 * Z2 * (ze + ze - 2 * a + A2 * iz) * rn_cond(M, x, y, Z, 11, L, gamma)
    rh = rh + 231 * A2 * AZ * Z2 * Z2 * Z * rn_cond(M, x, y, Z, 13, L, gamma)
    return (b * b * iz ** 3) * (0.375 * b * b / np.pi) * rh


# --------------------------------------------------------------------------
# Canonical model (likelihood.py:13-71)
# --------------------------------------------------------------------------
def _mn_params(p):
    return {k: p[k] for k in ('logM_disc', 'Rs_disc', 'Hs_disc', 'x_origin', 'y_origin',
                              'z_origin', 'dirx', 'diry', 'dirz')}


def density_func(x, y, z, params):
    """MN disc + T3 bar + V4 bulge density (likelihood.py:13-43)."""
    rho_mn = mn_density(x, y, z, _mn_params(params))
    M_bar = 10.0 ** params['logM_bar']
    rho_t3 = t3_density(x, y, z, M_bar, params['a_bar'], params['b_bar'], params['L_bar'], GAMMA_BAR)
    rho_v4 = v4_density(x, y, z, M_bar, V4_A, V4_B, V4_L, V4_GAMMA)
    return rho_mn + rho_t3 + rho_v4


def potential_func(x, y, z, params_baryon, params_halo):
    """NFW + MN + T3 + V4 potential (likelihood.py:46-71)."""
    phi_halo = nfw_potential(x, y, z, params_halo)
    phi_mn = mn_potential(x, y, z, _mn_params(params_baryon))
    M_bar = 10.0 ** params_baryon['logM_bar']
    phi_t3 = t3_potential(x, y, z, M_bar, params_baryon['a_bar'], params_baryon['b_bar'],
                          params_baryon['L_bar'], GAMMA_BAR)
    phi_v4 = v4_potential(x, y, z, M_bar, V4_A, V4_B, V4_L, V4_GAMMA)
    return phi_halo + phi_mn + phi_t3 + phi_v4


def accel(x, y, z, params_baryon, params_halo):
    """-grad Phi by forward-mode AD (model.py:152-157).  Returns (ax, ay, az, Phi)."""
    X, Y, Z = D.Dual.seed([x, y, z])
    phi = potential_func(X, Y, Z, params_baryon, params_halo)
    return -phi.g[0], -phi.g[1], -phi.g[2], phi.v
