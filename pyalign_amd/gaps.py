"""Gap cost models (the pyalign.gaps surface, pyalign/gaps.py:1-288).

A gap cost w(k) depends only on the gap length k.  Linear, affine and
constant-zero costs reduce to the solver's special cases through
``to_special_case()`` (gaps.py:13-14, 102-108, 140-143, 181-184); the MI355X
solver runs those with its O(n m) kernels.  The remaining models give
their ``costs(n)`` tables to the general Waterman-Smith-Beyer kernel
(O(n m (n + m)), solver.h:3047-3184).
Plotting helpers are presentation only and are not provided.
"""

import numpy as np


class GapCost:
    """A gap cost w(k) that depends only on the gap length k."""

    def to_tuple(self):
        raise NotImplementedError()

    def to_special_case(self):
        return {}

    def costs(self, n):
        raise NotImplementedError()

    @property
    def title(self):
        raise NotImplementedError()


class ConstantGapCost(GapCost):
    """w(k) = u for k > 0, w(0) = 0."""

    def __init__(self, u):
        assert u >= 0
        self._u = u

    def to_tuple(self):
        return "constant", self._u

    def to_special_case(self):
        # a zero constant is a zero linear cost
        return {"linear": 0} if self._u == 0 else {}

    def costs(self, n):
        w = np.full((n,), self._u, dtype=np.float32)
        w[0] = 0
        return w

    @property
    def title(self):
        return f"w(k) = {self._u:.1f}, w(0) = 0"


class LinearGapCost(GapCost):
    """w(k) = u k."""

    def __init__(self, u):
        assert u >= 0
        self._u = u

    def to_tuple(self):
        return "linear", self._u

    def to_special_case(self):
        return {"linear": self._u}

    def costs(self, n):
        return np.linspace(0.0, (n - 1) * self._u, n, dtype=np.float32)

    @property
    def title(self):
        return f"w(k) = {self._u:.1f} * k"


class AffineGapCost(GapCost):
    """w(k) = u + v k, given either as (open, extend) with u = open - extend,
    v = extend, or directly as (u=..., v=...)."""

    def __init__(self, open=None, extend=0, u=None, v=None):
        if open is not None:
            assert u is None and v is None
            if open < extend:
                raise RuntimeError(f"please ensure {open} = open >= extend = {extend}")
            self._u = open - extend
            self._v = extend
        else:
            assert extend == 0
            self._u = u
            self._v = v
        assert self._u >= 0
        assert self._v >= 0

    def to_tuple(self):
        return "affine", self._u, self._v

    def to_special_case(self):
        return {"affine": (self._u, self._v)}

    def costs(self, n):
        w = self._u + np.linspace(0.0, (n - 1) * self._v, n, dtype=np.float32)
        w[0] = 0
        return w

    @property
    def title(self):
        return f"w(k) = {self._u:.1f} + {self._v:.1f} * k, w(0) = 0"


class LogarithmicGapCost(GapCost):
    """w(k) = u + v ln(k)."""

    def __init__(self, u, v):
        assert u >= 0
        assert v >= 0
        self._u = u
        self._v = v

    def to_tuple(self):
        return "logarithmic", self._u, self._v

    def costs(self, n):
        assert n > 0
        w = np.full((n,), self._u, dtype=np.float32)
        w[1:] += self._v * np.log(np.arange(1, n), dtype=np.float32)
        w[0] = 0
        return w

    @property
    def title(self):
        return f"w(k) = {self._u:.1f} + {self._v:.1f} * ln(k), w(0) = 0"


class ExponentialGapCost(GapCost):
    """w(k) = 1 - u^(-k v)."""

    def __init__(self, u, v):
        assert u >= 0
        assert v >= 0
        self._u = u
        self._v = v

    def to_tuple(self):
        return "exponential", self._u, self._v

    def costs(self, n):
        # Python float arithmetic per entry, rounded once to float32
        # (gaps.py:244-248)
        return np.array([1 - (self._u ** -(k * self._v)) for k in range(n)], dtype=np.float32)

    @property
    def title(self):
        return f"w(k) = 1 - {self._u:.1f}^(-k * {self._v:.1f})"


class UserFuncGapCost(GapCost):
    """w(k) given by a Python function, w(0) = 0."""

    def __init__(self, costs_fn):
        self._costs_fn = costs_fn

    def to_tuple(self):
        return "user", self._costs_fn

    def costs(self, n):
        c = np.empty((n,), dtype=np.float32)
        c[0] = 0
        for k in range(1, n):
            c[k] = self._costs_fn(k)
        return c

    @property
    def title(self):
        return f"w(k) = {self._costs_fn}, w(0) = 0"


def smooth_gap_cost(k):
    """Exponentially saturating cost with w(k) = 0.5 at the given k (k > 0)."""
    if k > 0:
        return ExponentialGapCost(2, 1 / k)
    return ConstantGapCost(0)
