"""Host-side helpers of the model path (rotation matrix, NFW parameter maps).

makeRotationMatrix restates utils.py:164-179; logMenc_logc_to_logM_logRs and
its forward map restate utils.py:514-569.  These are scalar set-up math on the
host; the per-orbit work is all on the GPU.
"""
import math

import numpy as np

from .synth import logM_logRs_to_logMenc_logc  # noqa: F401  (re-export, utils.py:514-523)


def makeRotationMatrix(alpha, beta, gamma):
    """(Rz(gamma) Rx(beta) Rz(alpha))^T, angles in degrees (utils.py:164-179)."""
    a, b, g = np.radians(alpha), np.radians(beta), np.radians(gamma)

    def Rz(t):
        return np.array([[np.cos(t), -np.sin(t), 0.0], [np.sin(t), np.cos(t), 0.0], [0.0, 0.0, 1.0]])

    def Rx(t):
        return np.array([[1.0, 0.0, 0.0], [0.0, np.cos(t), -np.sin(t)], [0.0, np.sin(t), np.cos(t)]])
    return (Rz(g) @ Rx(b) @ Rz(a)).T


def logMenc_logc_to_logM_logRs(logM_enc, log_c, r_enc=10.0, Delta=200., rho_crit=277.54):
    """(logM_enc(<r_enc), log c) -> (logM, logRs) by six Newton steps (utils.py:526-569)."""
    M_enc = 10.0 ** logM_enc
    c = 10.0 ** log_c
    coeff = 4.0 * math.pi * Delta * rho_crit / 3.0
    ln10 = math.log(10.0)
    logM = logM_enc + 0.5
    for _ in range(6):
        M = 10.0 ** logM
        Rs = (M / coeff) ** (1.0 / 3.0) / c
        x = r_enc / Rs
        f_nfw = math.log(1.0 + x) - x / (1.0 + x)
        f = M * f_nfw - M_enc
        df = (f_nfw - x * x / (3.0 * (1.0 + x) ** 2)) * M * ln10
        logM = logM - f / df
    M_sol = 10.0 ** logM
    return logM, math.log10((M_sol / coeff) ** (1.0 / 3.0) / c)
