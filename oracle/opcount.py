"""Operation count of the canonical force evaluation (test/bench infrastructure).

Runs the oracle's potential (oracle/potential.py, the restatement of
likelihood.py:46-71) on tracing scalars that build the expression graph with
common subexpressions merged (as XLA's CSE merges the needle recurrences the
reference recomputes per order) and terms multiplied by a zero constant
dropped (the gamma = 0 / gamma = 1 halves of the needle convolution), then
counts the graph reachable from Phi.

F_acc, the flops of one (Phi, grad Phi) evaluation, follows the forward-mode
convention: every graph operation costs 1 for its value plus t(op) for each of
the 3 tangents: add/sub 1, mul 3, div 3, sqrt 2, log 2, atan 3, pow 2; with a
parameter constant, mul/div 1 and add/sub 0.  Parameter set-up (10**logM,
rotation matrices) is not counted.  F_STEP counts the BS3 attempt of
integrates.py:90-112 without its three force evaluations.
"""
import math

from . import potential as P

_TAN = {'add': 1, 'sub': 1, 'mul': 3, 'div': 3, 'sqrt': 2, 'log': 2, 'atan': 3, 'pow': 2,
        'cmul': 1, 'cdiv': 1, 'cadd': 0, 'rdiv': 1, 'neg': 0}


class Node:
    memo = {}
    _next = [0]

    def __init__(self, v, op=None, parents=()):
        self.v = float(v)
        self.op = op
        self.parents = parents
        self.id = Node._next[0]
        Node._next[0] += 1

    @classmethod
    def reset(cls):
        cls.memo = {}

    @staticmethod
    def make(op, key, v, parents):
        k = (op,) + key
        n = Node.memo.get(k)
        if n is None:
            n = Node(v, op, parents)
            Node.memo[k] = n
        return n

    def _bin(self, o, fn, op, cop, commut):
        if isinstance(o, Zero):
            return o._rbin(self, op)
        if isinstance(o, Node):
            a, b = (self, o) if not commut or self.id <= o.id else (o, self)
            return Node.make(op, (a.id, b.id), fn(self.v, o.v), (self, o))
        c = float(o)
        if cop == 'cmul' and c == 0.0:
            return ZERO
        if cop == 'cadd' and c == 0.0:
            return self
        if cop == 'cmul' and c == 1.0:
            return self
        return Node.make(cop, (self.id, c, op), fn(self.v, c), (self,))

    def __add__(self, o): return self._bin(o, lambda a, b: a + b, 'add', 'cadd', True)
    def __radd__(self, o): return self._bin(o, lambda a, b: b + a, 'add', 'cadd', True)
    def __sub__(self, o): return self._bin(o, lambda a, b: a - b, 'sub', 'cadd', False)

    def __rsub__(self, o):
        if isinstance(o, Node):
            return o - self
        return Node.make('cadd', (self.id, float(o), 'rsub'), float(o) - self.v, (self,))

    def __mul__(self, o): return self._bin(o, lambda a, b: a * b, 'mul', 'cmul', True)
    def __rmul__(self, o): return self._bin(o, lambda a, b: b * a, 'mul', 'cmul', True)
    def __truediv__(self, o): return self._bin(o, lambda a, b: a / b, 'div', 'cdiv', False)

    def __rtruediv__(self, o):
        return Node.make('rdiv', (self.id, float(o)), float(o) / self.v, (self,))

    def __neg__(self):
        return Node.make('neg', (self.id,), -self.v, (self,))

    def __pow__(self, p):
        p = float(p)
        if p == int(p) and p >= 1:
            r = self
            for _ in range(int(p) - 1):
                r = r * self
            return r
        if p == 0.5:
            return self.sqrt()
        return Node.make('pow', (self.id, p), self.v ** p, (self,))

    def sqrt(self):
        return Node.make('sqrt', (self.id,), math.sqrt(self.v), (self,))

    def log(self):
        return Node.make('log', (self.id,), math.log(self.v), (self,))

    def arctan(self):
        return Node.make('atan', (self.id,), math.atan(self.v), (self,))


class Zero:
    """A term multiplied by an exact zero constant."""

    def _rbin(self, other, op):
        if op in ('add',):
            return other
        if op == 'sub':
            return other
        return self

    def __add__(self, o): return o
    __radd__ = __add__

    def __sub__(self, o): return -o if isinstance(o, Node) else -float(o)
    def __rsub__(self, o): return o
    def __mul__(self, o): return self
    __rmul__ = __mul__
    def __truediv__(self, o): return self
    def __neg__(self): return self


ZERO = Zero()

# BS3 attempt without its force evaluations (integrates.py:90-112): two stage
# states 2*(1 + 6*2), three RHS assemblies 3*(4*2), y_new 6*(2+2+2+1+1),
# error 6*(2+2+2+1+1) and its scaling 6*(2 + 1 + 2) + 2, controller 6.
F_STEP = 2 * 13 + 3 * 8 + 6 * 8 + 6 * 8 + 6 * 5 + 2 + 6


def force_flops(params_baryon, params_halo, point=(3.0, 1.0, 0.3)):
    """(F_value, F_acc, tally) for one evaluation of Phi / (Phi, grad Phi)."""
    Node.reset()
    x, y, z = (Node(v, 'input') for v in point)
    out = P.potential_func(x, y, z, params_baryon, params_halo)
    seen, stack, tally = set(), [out], {}
    while stack:
        n = stack.pop()
        if not isinstance(n, Node) or n.id in seen or n.op == 'input':
            continue
        seen.add(n.id)
        tally[n.op] = tally.get(n.op, 0) + 1
        stack.extend(n.parents)
    f_val = sum(tally.values())
    f_acc = sum(c * (1 + 3 * _TAN[op]) for op, c in tally.items())
    return f_val, f_acc, tally
