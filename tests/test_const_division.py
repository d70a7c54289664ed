"""The Stumpff series of the GPU Kepler drift divides by the constants 6,
12, 20, ..., 210 with one multiply and two FMAs (q = RN(z y) with
y = RN(1/b), r = fma(-q, b, z), RN(fma(r, y, q))) instead of a full
division.  By Markstein's theorem this is the correctly rounded z / b for
normal quotients; here the sequence is evaluated with exact rational
arithmetic (each FMA rounded once) on random and boundary z and compared
with IEEE division.  (CPU only; the bitwise WHFast GPU tests check the same
kernels end to end.)"""
from fractions import Fraction

import numpy as np
import pytest

CONSTS = [6.0, 12.0, 20.0, 30.0, 42.0, 56.0, 72.0, 90.0, 110.0, 132.0, 156.0, 182.0, 210.0]


def _fma(a, b, c):
    # one rounding of the exact a * b + c (Fraction -> float rounds to nearest even)
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def _div_const(z, b):
    y = 1.0 / b
    q = z * y
    r = _fma(-q, b, z)
    assert Fraction(r) == Fraction(z) - Fraction(b) * Fraction(q)  # the remainder is exact
    return _fma(r, y, q)


@pytest.mark.parametrize("b", CONSTS)
def test_fma_division_by_constant_is_correctly_rounded(b):
    rng = np.random.default_rng(int(b))
    mant = rng.integers(0, 1 << 52, 1500, dtype=np.int64)
    exps = rng.integers(-899, 2, 1500)
    zs = [np.ldexp(1.0 + float(mm) / 2.0**52, int(e)) * (1 if k % 2 else -1) for k, (mm, e) in
          enumerate(zip(mant, exps))]
    # quotients next to powers of two and to small odd multiples
    for e in (-40, -7, 0):
        for mult in (1, 3, 5, 7):
            z0 = np.ldexp(b * mult, e)
            zs += [float(np.nextafter(z0, np.inf) if j > 0 else np.nextafter(z0, -np.inf)) for j in (-1, 1)]
            zs.append(z0)
    for z in zs:
        assert _div_const(z, b) == z / b, (b, z.hex())
