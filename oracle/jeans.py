"""Axisymmetric Jeans initial conditions, rotation curve and orbital time.

Restates jeans.py:9-94 (Jeans moments, trapezoid rule on 1000 points,
finite-difference derivatives with d = 5e-3), utils.py:144-162 (clockwise
cylindrical -> Cartesian) and utils.py:182-273 (rotation curve and orbital
timescale by central differences with dR = 1e-3).  The JAX normals drawn from
keys 42/109/2026 (jeans.py:89) are an explicit input ``normals`` (n, 3).
"""
import numpy as np


def _trapezoid(yv, xv):
    """jax.scipy.integrate.trapezoid with sample points (last axis)."""
    return 0.5 * np.sum(np.diff(xv, axis=-1) * (yv[..., 1:] + yv[..., :-1]), axis=-1)


def jeans_moments(xs, ys, zs, pb, ph, potential_func, density_func, b=1.0, npts=1000):
    """get_jeans_moments (jeans.py:9-79), vectorised over stars."""
    xs, ys, zs = (np.asarray(a, dtype=np.float64) for a in (xs, ys, zs))
    R_star = np.sqrt(xs ** 2 + ys ** 2)
    d = 5e-3

    def dphi_dz(x, y, z):
        return (potential_func(x, y, z + d, pb, ph) - potential_func(x, y, z - d, pb, ph)) / (2 * d)

    def dphi_dR(x, y, z):
        R = np.sqrt(x ** 2 + y ** 2)
        return (potential_func(R + d, 0, z, pb, ph) - potential_func(R - d, 0, z, pb, ph)) / (2 * d)

    pts = np.linspace(np.abs(zs), 10.0, npts, axis=-1)          # (n, npts)
    X = xs[:, None]
    Y = ys[:, None]
    integ = density_func(X, Y, pts, pb) * dphi_dz(X, Y, pts)
    integral = _trapezoid(integ, pts)
    nu = density_func(xs, ys, zs, pb)
    with np.errstate(divide='ignore', invalid='ignore'):
        sz2 = np.maximum((1.0 / nu) * integral, 0.0)
    sz = np.sqrt(sz2)
    sR2 = b * sz2
    sR = np.sqrt(sR2)

    def vertical_pressure(r_in):
        rr = r_in[:, None]
        v = density_func(rr, 0.0, pts, pb) * dphi_dz(rr, 0.0, pts)
        return _trapezoid(v, pts)

    dR = 5e-3
    dnu = b * (vertical_pressure(R_star + dR) - vertical_pressure(R_star - dR)) / (2 * dR)
    with np.errstate(divide='ignore', invalid='ignore'):
        vtot2 = sR2 + (R_star / nu) * dnu + R_star * dphi_dR(xs, ys, zs)
    sphi = sR
    vmean = np.sqrt(np.maximum(vtot2 - sphi ** 2, 0.0))
    bad = nu <= 0
    z0 = np.zeros_like(vmean)
    return (np.where(bad, z0, vmean), np.where(bad, z0, sR),
            np.where(bad, z0, sz), np.where(bad, z0, sphi))


def cyl_to_cart_clockwise(R, phi, vR, vphi):
    """utils.py:144-162."""
    c, s = np.cos(phi), np.sin(phi)
    return R * c, R * s, vR * c + vphi * s, vR * s - vphi * c


def get_w0_new(w0, normals, pb, ph, potential_func, density_func):
    """jeans.py:83-94 with the three JAX normal draws given as normals (n, 3)."""
    w0 = np.asarray(w0, dtype=np.float64)
    v_rot, sR, sz, sphi = jeans_moments(w0[:, 0], w0[:, 1], w0[:, 2], pb, ph,
                                        potential_func, density_func)
    g1, g2, g3 = normals[:, 0], normals[:, 1], normals[:, 2]
    vR = g1 * sR
    vz = g2 * sz
    vphi = v_rot + g3 * sphi
    x, y, vx, vy = cyl_to_cart_clockwise(np.sqrt(w0[:, 0] ** 2 + w0[:, 1] ** 2),
                                         np.arctan2(w0[:, 1], w0[:, 0]), vR, vphi)
    return np.stack([x, y, w0[:, 2], vx, vy, vz], axis=-1)


def _dphi_dR_fd(R, z, potential_func, pb, ph, dR):
    return (potential_func(R + dR, 0.0, z, pb, ph) - potential_func(R - dR, 0.0, z, pb, ph)) / (2.0 * dR)


def rotation_curve(R, z, potential_func, pb, ph, dR=1e-3):
    """get_rotation_curve (utils.py:231-273)."""
    R_safe = np.maximum(np.abs(R), 2 * dR)
    d = _dphi_dR_fd(R_safe, z, potential_func, pb, ph, dR)
    return np.sqrt(np.maximum(R_safe * d, 0.0))


def orbital_timescale(R, z, potential_func, pb, ph, dR=1e-3):
    """estimate_orbital_timescale (utils.py:182-228)."""
    R_safe = np.maximum(np.abs(R), 2 * dR)
    d = _dphi_dR_fd(R_safe, z, potential_func, pb, ph, dR)
    omega = np.sqrt(np.maximum(np.abs(d) / R_safe, 1e-20))
    return 2 * np.pi / omega
