"""Julia Float64 range semantics for the mesh and particle lattices.

The reference builds node and particle coordinates with ``collect(a:s:b)``
(``get_coords.jl:22-55``, ``mpts_populate.jl:29-43``) and classifies B-spline
nodes by floating-point equality with those values (``bsmpm.jl:63-76``).
Julia does not compute ``a + (i-1)*s`` for such a range: ``Base.(:)`` first
looks for small rationals that round to ``a``, ``s`` and ``b`` (``Base.rat``)
and, when it finds them, returns every element as the correctly rounded value
of ``(a_n + (i-1)*s_n)/den`` (a ``TwicePrecision`` range).  Otherwise it
falls back to ``a + (i-1)*s`` with one rounding of the product.  This module
restates that algorithm (Julia 1.10 ``base/twiceprecision.jl``) so the
host-side mesh holds the same coordinates the reference would hold.
"""
from __future__ import annotations

import math
from fractions import Fraction

import numpy as np

_MAXINT_F32 = 16777216          # maxintfloat(Float32, Int)
_MAXINT_F64 = 9007199254740992  # maxintfloat(Float64, Int)
_I64_MIN, _I64_MAX = -(1 << 63), (1 << 63) - 1


def _wrap64(v: int) -> int:
    """Int64 two's-complement wrap, as Julia's integer arithmetic does."""
    v &= (1 << 64) - 1
    return v - (1 << 64) if v >= (1 << 63) else v


def rat(x: float) -> tuple[int, int]:
    """``Base.rat``: continued-fraction rational approximation in floats."""
    y = x
    a = d = 1
    b = c = 0
    m = _MAXINT_F32
    while abs(y) <= m:
        f = int(math.trunc(y))
        y -= f
        a, c = f * a + c, a
        b, d = f * b + d, b
        if not max(abs(a), abs(b)) <= m:
            return c, d
        if float(a) / float(b) == x:
            break
        y = 1.0 / y if y != 0.0 else math.inf
    return a, b


def _isbetween(a: float, x: float, b: float) -> bool:
    return (a <= x <= b) or (b <= x <= a)


def _jround_int(x: float) -> int:
    """``round(Int, x)``: round half to even."""
    return int(round(x))


def frange(start: float, step: float, stop: float) -> np.ndarray:
    """``collect(start:step:stop)`` for Float64 arguments."""
    start, step, stop = float(start), float(step), float(stop)
    if step == 0.0:
        raise ValueError("range step cannot be zero")
    step_n, step_d = rat(step)
    if step_d != 0 and float(step_n) / float(step_d) == step:
        start_n, start_d = rat(start)
        stop_n, stop_d = rat(stop)
        if (start_d != 0 and stop_d != 0 and float(start_n) / float(start_d) == start
                and float(stop_n) / float(stop_d) == stop):
            den = _wrap64(start_d * step_d // math.gcd(start_d, step_d)) if start_d and step_d else 0
            m = _MAXINT_F64
            if (den != 0 and abs(start * den) <= m and abs(step * den) <= m
                    and den % start_d == 0 and den % step_d == 0):
                start_n = _jround_int(start * den)
                step_n = _jround_int(step * den)
                num = _wrap64(_wrap64(den * stop_n) - _wrap64(stop_d * start_n) + _wrap64(step_n * stop_d))
                dv = _wrap64(step_n * stop_d)
                q = abs(num) // abs(dv)
                q = q if (num >= 0) == (dv >= 0) else -q        # Int div truncates
                ln = max(0, q)
                if (_isbetween(start, start + (ln - 1) * step, stop + step / 2.0)
                        and not _isbetween(start, start + ln * step, stop)):
                    return np.array([float(Fraction(start_n + i * step_n, den)) for i in range(ln)],
                                    dtype=np.float64)
    lf = (stop - start) / step
    if lf < 0:
        ln = 0
    elif lf == 0:
        ln = 1
    else:
        ln = _jround_int(lf) + 1
        stop2 = start + (ln - 1) * step
        ln -= int(start < stop < stop2) + int(start > stop > stop2)
    # steprangelen_hp(Float64, start, step, 0, len, 1): ref + u*step with the
    # product rounded once and the sum rounded once.
    return np.array([start + (i * step) for i in range(ln)], dtype=np.float64)


def linrange(a: float, b: float, n: int) -> np.ndarray:
    """``collect(LinRange(a, b, n))`` (``Base.lerpi``)."""
    a, b = float(a), float(b)
    if n == 1:
        return np.array([a])
    d = n - 1
    out = np.empty(n)
    for j in range(n):
        t = j / d
        out[j] = (1.0 - t) * a + t * b
    return out
