"""Potential parameter packing for the device kernels.

Mirrors the parameter dicts of potentials.py:18-65 (NFW, Miyamoto-Nagai) and
the T3 / V4 arguments built by likelihood.py:46-71, and packs them into the
``smx_potential`` struct the HIP kernels read.  Derived constants are formed
with the same operations the reference applies (``-G * 10**logM``,
``Hs**2 * 10**logM / (4 pi)``).
"""
import math

import numpy as np

from . import _lib

G = 1.3271244e20 * 3.15576e16 ** 2 / 3.0856775814913673e19 ** 3   # constants.py:5
EPSILON = 1e-12                                                     # constants.py:11
KPCGYR_TO_KMS = 3.0856775814913673e16 / 3.15576e16                  # constants.py:10
V4_A, V4_B, V4_L, V4_GAMMA = 0.5, 0.5, 0.1, 0.0                     # likelihood.py:10
GAMMA_BAR = 1.0                                                     # likelihood.py:11


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


def _f(v):
    return float(np.asarray(v))


def pack_potential(params_baryon, params_halo=None):
    """smx_potential from the reference's params_baryon / params_halo dicts.

    ``params_baryon`` needs the Miyamoto-Nagai and bar keys of model.py:87-102;
    ``params_halo`` the NFW keys of likelihood.py:132-144 (omit it for density
    only evaluations).
    """
    p = _lib.SmxPotential()
    if params_halo is not None:
        ph = params_halo
        p.nfw_mGM = -G * 10.0 ** _f(ph['logM'])
        p.nfw_Rs = _f(ph['Rs'])
        p.nfw_a, p.nfw_b, p.nfw_c = _f(ph['a']), _f(ph['b']), _f(ph['c'])
        p.nfw_origin[:] = [_f(ph['x_origin']), _f(ph['y_origin']), _f(ph['z_origin'])]
        p.nfw_rot[:] = list(get_mat(_f(ph['dirx']), _f(ph['diry']), _f(ph['dirz'])).ravel())
    else:
        p.nfw_mGM, p.nfw_Rs, p.nfw_a, p.nfw_b, p.nfw_c = 0.0, 1.0, 1.0, 1.0, 1.0
        p.nfw_rot[:] = list(np.eye(3).ravel())
    pb = params_baryon
    M_disc = 10.0 ** _f(pb['logM_disc'])
    p.mn_mGM = -G * M_disc
    p.mn_rho0 = _f(pb['Hs_disc']) ** 2 * M_disc / (4 * math.pi)
    p.mn_Rs, p.mn_Hs = _f(pb['Rs_disc']), _f(pb['Hs_disc'])
    p.mn_origin[:] = [_f(pb['x_origin']), _f(pb['y_origin']), _f(pb['z_origin'])]
    p.mn_rot[:] = list(get_mat(_f(pb['dirx']), _f(pb['diry']), _f(pb['dirz'])).ravel())
    M_bar = 10.0 ** _f(pb['logM_bar'])
    p.t3_M, p.t3_a, p.t3_b = M_bar, _f(pb['a_bar']), _f(pb['b_bar'])
    p.t3_L, p.t3_gamma = _f(pb['L_bar']), GAMMA_BAR
    p.v4_M, p.v4_a, p.v4_b, p.v4_L, p.v4_gamma = M_bar, V4_A, V4_B, V4_L, V4_GAMMA
    p.G = G
    return p


def _points(x, y, z):
    import torch
    x, y, z = (_lib.dev(a) for a in (x, y, z))
    x, y, z = torch.broadcast_tensors(x, y, z)
    return torch.stack([x, y, z], dim=-1).contiguous(), x.shape


def potential_eval(x, y, z, params_baryon, params_halo, with_accel=False):
    """Phi (and -grad Phi) at points, on the GPU."""
    import torch
    L = _lib.lib()
    xyz, shape = _points(x, y, z)
    n = xyz.numel() // 3
    phi = torch.empty(shape, dtype=torch.float64, device='cuda')
    acc = torch.empty(shape + (3,), dtype=torch.float64, device='cuda') if with_accel else None
    pot = pack_potential(params_baryon, params_halo)
    _lib.check(L.smx_potential_eval(pot, _lib.ptr(xyz), n, _lib.ptr(phi), _lib.ptr(acc),
                                    _lib.stream_ptr()), 'smx_potential_eval')
    return (phi, acc) if with_accel else phi


def density_eval(x, y, z, params):
    """Stellar density MN + T3 + V4 at points, on the GPU."""
    import torch
    L = _lib.lib()
    xyz, shape = _points(x, y, z)
    n = xyz.numel() // 3
    rho = torch.empty(shape, dtype=torch.float64, device='cuda')
    _lib.check(L.smx_density_eval(pack_potential(params), _lib.ptr(xyz), n, _lib.ptr(rho),
                                  _lib.stream_ptr()), 'smx_density_eval')
    return rho
