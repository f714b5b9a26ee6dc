"""Exact-arithmetic brute force used to pin the oracle on tiny inputs.

Floating-point operations are evaluated in exact rational arithmetic and
rounded once to the nearest double (CPython's int/int true division is
correctly rounded), so fma(a, b, c) = round(a*b + c) and a + b = round(a + b)
are computed independently of any compiler or evaluation order.
"""
from __future__ import annotations

from fractions import Fraction

import numpy as np


def rnd(q: Fraction) -> float:
    return q.numerator / q.denominator


def fma(a: float, b: float, c: float) -> float:
    return rnd(Fraction(a) * Fraction(b) + Fraction(c))


def add(a: float, b: float) -> float:
    return rnd(Fraction(a) + Fraction(b))


def chain(arow, bcol) -> float:
    """Sequential fma chain from +0.0, the reference kernel's per-element order."""
    acc = 0.0
    for a, b in zip(arow, bcol):
        acc = fma(float(a), float(b), acc)
    return acc


def exact_dot(arow, bcol) -> Fraction:
    return sum((Fraction(float(a)) * Fraction(float(b)) for a, b in zip(arow, bcol)), Fraction(0))


def exact_product(A: np.ndarray, B: np.ndarray):
    m, k = A.shape
    n = B.shape[1]
    return [[exact_dot(A[i, :], B[:, j]) for j in range(n)] for i in range(m)]


def rounded_exact_product(A: np.ndarray, B: np.ndarray) -> np.ndarray:
    P = exact_product(A, B)
    return np.array([[rnd(x) for x in row] for row in P], dtype=np.float64)
