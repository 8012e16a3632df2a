"""Gauss-Hermite helpers on device tensors (ghmoments.py:53-79)."""
import math

import torch


def h_to_V_sigma(h1, h2, v0, s, h3=0.0, h4=0.0):
    """GH coefficients -> LOSVD mean and dispersion (ghmoments.py:53-79)."""
    sqrt2, sqrt3, sqrt6 = math.sqrt(2.0), math.sqrt(3.0), math.sqrt(6.0)
    delta_v = sqrt2 * h1 + sqrt3 * h3
    v_mean = v0 + s * delta_v
    sig2 = 1.0 + 2.0 * sqrt2 * h2 + 2.0 * sqrt6 * h4 - delta_v ** 2
    return v_mean, s * torch.sqrt(torch.clamp(sig2, min=1e-10))
