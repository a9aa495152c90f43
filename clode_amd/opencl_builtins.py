"""Python versions of the OpenCL built-ins a right-hand side may call
(counterpart of clode/opencl_builtins.py), so that a Python-defined
right-hand side runs as Python (for checking and type checking) and converts
to the device program unchanged.  Functions that exist in ``math`` are
re-exported; the others follow their OpenCL definitions."""
from __future__ import annotations

from math import (acos, acosh, asin, asinh, atan, atan2, atanh, ceil, copysign, cos, cosh, erf, erfc, exp,
                  expm1, fabs, floor, fmod, gamma, hypot, ldexp, lgamma, log, log1p, log2, log10, pi, pow, sin,
                  sinh, sqrt, tan, tanh, trunc)
from math import nextafter as _nextafter
from math import remainder

import numpy as _np


def acospi(x: float) -> float:
    return acos(x) / pi


def asinpi(x: float) -> float:
    return asin(x) / pi


def atanpi(x: float) -> float:
    return atan(x) / pi


def atan2pi(y: float, x: float) -> float:
    return atan2(y, x) / pi


def cbrt(x: float) -> float:
    return float(_np.cbrt(x))


def cospi(x: float) -> float:
    return cos(pi * x)


def sinpi(x: float) -> float:
    return sin(pi * x)


def tanpi(x: float) -> float:
    return tan(pi * x)


def exp2(x: float) -> float:
    return 2.0 ** x


def exp10(x: float) -> float:
    return 10.0 ** x


def fdim(x: float, y: float) -> float:
    return x - y if x > y else 0.0


def heaviside(x: float) -> float:
    return 1.0 if x >= 0.0 else 0.0


def ilogb(x: float) -> int:
    return int(floor(log2(abs(x))))


def nextafter(x: float, y: float) -> float:
    return _nextafter(x, y)


def pown(x: float, n: int) -> float:
    return x ** n


def powr(x: float, y: float) -> float:
    return x ** y


def rint(x: float) -> float:
    return float(_np.rint(x))


def rootn(x: float, n: int) -> float:
    return float(_np.sign(x) * abs(x) ** (1.0 / n)) if n % 2 else x ** (1.0 / n)


def rsqrt(x: float) -> float:
    return 1.0 / sqrt(x)


__all__ = ["acos", "acosh", "acospi", "asin", "asinh", "asinpi", "atan", "atan2", "atan2pi", "atanh", "atanpi",
           "cbrt", "ceil", "copysign", "cos", "cosh", "cospi", "erf", "erfc", "exp", "exp2", "exp10", "expm1",
           "fabs", "fdim", "floor", "fmod", "gamma", "heaviside", "hypot", "ilogb", "ldexp", "lgamma", "log",
           "log1p", "log2", "log10", "nextafter", "pi", "pow", "pown", "powr", "remainder", "rint", "rootn",
           "rsqrt", "sin", "sinh", "sinpi", "sqrt", "tan", "tanh", "tanpi", "trunc"]
