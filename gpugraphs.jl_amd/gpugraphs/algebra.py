"""Semiring operators of src/algebra.jl.

Each ``GPUGraphs_*`` operator is a callable with the reference's 6-argument
signature ``(x, y, A_i, A_j, B_i, B_j)`` (src/algebra.jl:28-111) that also
carries the operator code the C ABI dispatches on.  Passing anything else as
``mul`` / ``add`` / ``accum`` raises ``MethodError``: the kernels are compiled
for these operators, not for arbitrary closures.
"""
from __future__ import annotations

import numpy as np

from ._lib import ArgumentError, MethodError


class Operator:
    __slots__ = ("name", "code", "_f")

    def __init__(self, name: str, code: int, f):
        self.name, self.code, self._f = name, code, f

    def __call__(self, x, y, A_i=None, A_j=None, B_i=None, B_j=None):
        return self._f(x, y, A_i, A_j, B_i, B_j)

    def __repr__(self):
        return self.name


def _max(x, y):
    if isinstance(x, float) and x != x:
        return x
    return x if x > y else y


def _min(x, y):
    if isinstance(x, float) and x != x:
        return x
    return x if x < y else y


def _typemax(v):
    t = np.asarray(v).dtype
    if t == np.bool_:
        return True
    if np.issubdtype(t, np.integer):
        return np.iinfo(t).max
    return np.inf


def _safe_add(x, y, *_):
    # src/algebra.jl:47-48: saturate at typemax
    sat = max(_typemax(x), _typemax(y))
    return sat if _max(x, y) == sat else x + y


GPUGraphs_first = Operator("GPUGraphs_first", 1, lambda x, y, *_: x)
GPUGraphs_second = Operator("GPUGraphs_second", 2, lambda x, y, *_: y)
GPUGraphs_any = Operator("GPUGraphs_any", 3, lambda x, y, *_: _max(x, y))
GPUGraphs_pair = Operator("GPUGraphs_pair", 4, lambda x, y, *_: type(x)(1) if y != 0 else type(x)(0))
GPUGraphs_add = Operator("GPUGraphs_add", 5, lambda x, y, *_: x + y)
GPUGraphs_safe_add = Operator("GPUGraphs_safe_add", 6, _safe_add)
GPUGraphs_mul = Operator("GPUGraphs_mul", 7, lambda x, y, *_: x * y)
GPUGraphs_min = Operator("GPUGraphs_min", 8, lambda x, y, *_: _min(x, y))
GPUGraphs_max = Operator("GPUGraphs_max", 9, lambda x, y, *_: _max(x, y))
GPUGraphs_lor = Operator("GPUGraphs_lor", 10, lambda x, y, *_: x | y)
GPUGraphs_land = Operator("GPUGraphs_land", 11, lambda x, y, *_: x & y)
GPUGraphs_bor = Operator("GPUGraphs_bor", 12, lambda x, y, *_: x | y)
GPUGraphs_band = Operator("GPUGraphs_band", 13, lambda x, y, *_: x & y)
GPUGraphs_secondi = Operator("GPUGraphs_secondi", 14, lambda x, y, A_i, A_j, B_i, B_j: B_i * y)

OPERATORS = [GPUGraphs_first, GPUGraphs_second, GPUGraphs_any, GPUGraphs_pair, GPUGraphs_add,
             GPUGraphs_safe_add, GPUGraphs_mul, GPUGraphs_min, GPUGraphs_max, GPUGraphs_lor,
             GPUGraphs_land, GPUGraphs_bor, GPUGraphs_band, GPUGraphs_secondi]


def op_code(op) -> int:
    if isinstance(op, Operator):
        return op.code
    raise MethodError(f"{op!r} is not one of the GPUGraphs_* operators with a HIP kernel")


def _np_type(T):
    return np.dtype(T)


def monoid_neutral(T, op):
    """Neutral element of ``op`` over type ``T`` (src/algebra.jl:10-23,36,44,...)."""
    t = _np_type(T)
    info = (np.iinfo(t) if np.issubdtype(t, np.integer) else None)
    if op in (GPUGraphs_add, GPUGraphs_any, GPUGraphs_bor, GPUGraphs_safe_add):
        return t.type(0)
    if op is GPUGraphs_mul:
        return t.type(1)
    if op is GPUGraphs_min:
        return True if t == np.bool_ else (t.type(info.max) if info else t.type(np.inf))
    if op is GPUGraphs_max:
        return False if t == np.bool_ else (t.type(info.min) if info else t.type(-np.inf))
    if op is GPUGraphs_band:
        return True if t == np.bool_ else t.type(~0)
    if op is GPUGraphs_lor and t == np.bool_:
        return False
    if op is GPUGraphs_land and t == np.bool_:
        return True
    raise ArgumentError(f"monoid_neutral not implemented for type {t} and operation {op}")


def monoid_absorb(T, op):
    """Absorbing element, or None when the monoid has none (src/algebra.jl:14,37,...)."""
    t = _np_type(T)
    info = (np.iinfo(t) if np.issubdtype(t, np.integer) else None)
    if op in (GPUGraphs_add, GPUGraphs_safe_add, GPUGraphs_max):
        return True if t == np.bool_ else (t.type(info.max) if info else t.type(np.inf))
    if op is GPUGraphs_min:
        return False if t == np.bool_ else (t.type(info.min) if info else t.type(-np.inf))
    if op is GPUGraphs_mul or op is GPUGraphs_band:
        return t.type(0)
    if op is GPUGraphs_bor:
        return True if t == np.bool_ else t.type(~0)
    if op is GPUGraphs_lor and t == np.bool_:
        return True
    if op is GPUGraphs_land and t == np.bool_:
        return False
    return None  # GPUGraphs_any's absorb is the type Any: never equal to a value
