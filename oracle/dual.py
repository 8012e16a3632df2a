"""Vectorised forward-mode dual numbers (value + k tangents) over NumPy arrays.

The reference gets forces as ``-jax.grad(potential)`` (model.py:152-157,
dehnen_bar.py:1546-1549): the exact derivative of the same expression graph.
This class gives the same thing on the CPU: every elementwise operation carries
its tangents, so ``Dual.seed([x, y, z])`` pushed through the potential returns
Phi and grad Phi together.  Pinned against central finite differences in
tests/test_oracle_potential.py.
"""
import numpy as np


class Dual:
    __slots__ = ("v", "g")

    def __init__(self, v, g):
        self.v = v          # ndarray
        self.g = g          # ndarray, shape (k,) + v.shape

    @staticmethod
    def seed(values):
        """Independent variables: values[i] gets unit tangent e_i."""
        vals = [np.asarray(v, dtype=np.float64) for v in values]
        shape = np.broadcast_shapes(*[v.shape for v in vals])
        k = len(vals)
        out = []
        for i, v in enumerate(vals):
            g = np.zeros((k,) + shape)
            g[i] = 1.0
            out.append(Dual(np.broadcast_to(v, shape).copy(), g))
        return out

    # -- helpers -----------------------------------------------------------
    @staticmethod
    def _lift(o, like):
        if isinstance(o, Dual):
            return o
        o = np.asarray(o, dtype=np.float64)
        return Dual(o, np.zeros((like.g.shape[0],) + np.broadcast_shapes(o.shape, like.v.shape)))

    # -- arithmetic --------------------------------------------------------
    def __neg__(self):
        return Dual(-self.v, -self.g)

    def __add__(self, o):
        if not isinstance(o, Dual):
            return Dual(self.v + o, self.g + np.zeros_like(np.asarray(o, dtype=np.float64)))
        return Dual(self.v + o.v, self.g + o.g)

    __radd__ = __add__

    def __sub__(self, o):
        if not isinstance(o, Dual):
            return Dual(self.v - o, self.g + np.zeros_like(np.asarray(o, dtype=np.float64)))
        return Dual(self.v - o.v, self.g - o.g)

    def __rsub__(self, o):
        return Dual(o - self.v, -self.g + np.zeros_like(np.asarray(o, dtype=np.float64)))

    def __mul__(self, o):
        if not isinstance(o, Dual):
            return Dual(self.v * o, self.g * o)
        return Dual(self.v * o.v, self.g * o.v + self.v * o.g)

    __rmul__ = __mul__

    def __truediv__(self, o):
        if not isinstance(o, Dual):
            return Dual(self.v / o, self.g / o)
        q = self.v / o.v
        return Dual(q, (self.g - q * o.g) / o.v)

    def __rtruediv__(self, o):
        q = o / self.v
        return Dual(q, -q * self.g / self.v)

    def __pow__(self, p):
        # only real scalar exponents appear on the path
        p = float(p)
        if p == 2.0:
            return self * self
        if p == 3.0:
            return self * self * self
        vp = np.power(self.v, p)
        return Dual(vp, p * np.power(self.v, p - 1.0) * self.g)


def sqrt(a):
    if isinstance(a, Dual):
        s = np.sqrt(a.v)
        return Dual(s, a.g * (0.5 / s))
    return np.sqrt(a)


def log(a):
    if isinstance(a, Dual):
        return Dual(np.log(a.v), a.g / a.v)
    return np.log(a)


def arctan(a):
    if isinstance(a, Dual):
        return Dual(np.arctan(a.v), a.g / (1.0 + a.v * a.v))
    return np.arctan(a)


def value(a):
    return a.v if isinstance(a, Dual) else a
