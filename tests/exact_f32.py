"""Exact binary32 arithmetic for tests: every operation is computed exactly in
rationals and rounded once to the nearest float32 (ties to even), so fused
multiply-adds and plain products can be told apart without relying on the
platform's float or fma.  Normal range only (no subnormals, no overflow)."""
from fractions import Fraction

import numpy as np


def rne(q):
    """Nearest float32 to the rational q (round half to even)."""
    q = Fraction(q)
    if q == 0:
        return np.float32(0.0)
    sign = -1 if q < 0 else 1
    a = abs(q)
    e = a.numerator.bit_length() - a.denominator.bit_length()
    if Fraction(2) ** e > a:
        e -= 1
    # a in [2^e, 2^(e+1)): 24 significant bits -> unit 2^(e-23)
    scaled = a / Fraction(2) ** (e - 23)
    m = scaled.numerator // scaled.denominator
    rem = scaled - m
    if rem > Fraction(1, 2) or (rem == Fraction(1, 2) and m % 2 == 1):
        m += 1
    return np.float32(sign * float(Fraction(m) * Fraction(2) ** (e - 23)))


def F(x):
    return Fraction(float(np.float32(x)))


def mul(a, b):
    return rne(F(a) * F(b))


def add(a, b):
    return rne(F(a) + F(b))


def sub(a, b):
    return rne(F(a) - F(b))


def div(a, b):
    return rne(F(a) / F(b))


def fma(a, b, c):
    return rne(F(a) * F(b) + F(c))
