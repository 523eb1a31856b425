"""The oracle's double-double sin/cos (the same arithmetic as the device's
relaxation angles) against exact values: pi from Machin's formula and the Taylor
series in exact rational arithmetic, rounded once to double.  Every result must be
the correctly rounded one, including arguments next to multiples of pi/2 where the
reduction cancels most bits."""
import math
from fractions import Fraction as F

import numpy as np
import pytest


def _arctan_inv(x, bits):
    one = 1 << bits
    s, t, n, sign = 0, one // x, 1, 1
    while t:
        s += sign * (t // n)
        t //= x * x
        n += 2
        sign = -sign
    return F(s, one)


PI = 16 * _arctan_inv(5, 320) - 4 * _arctan_inv(239, 320)


def exact_sincos(x):
    X = F(x)
    k = round(X / (PI / 2))
    r = X - k * PI / 2

    def series(t, first):
        acc, n = F(0), first
        while t != 0 and (acc == 0 or abs(t) > abs(acc) / 2 ** 200):
            acc += t
            t = -t * r * r / ((n + 1) * (n + 2))
            n += 2
        return acc

    s, c = series(r, 1), series(F(1), 0)
    S, C = [(s, c), (c, -s), (-s, -c), (-c, s)][k % 4]
    return float(S), float(C)       # Fraction -> float rounds correctly


def _args():
    rng = np.random.default_rng(7)
    xs = [float(v) for v in rng.random(500) * 2 * np.pi] + [float(v) for v in rng.random(200) * np.pi / 2]
    for k in range(1, 9):
        b = float(k * PI / 2)
        xs += [b, math.nextafter(b, 0.0), math.nextafter(b, 10.0), b * (1 + 1e-12)]
    # the second reduction r = j pi/128 + t: t near 0 and near +-pi/256 (the table
    # switches to the next entry), for j across the range
    for j in range(-32, 33, 3):
        for q in (0, 4):
            b = float(q * PI / 2 + j * PI / 128)
            h = float(PI / 256)
            xs += [b, math.nextafter(b, 0.0), math.nextafter(b, 10.0), b + 0.999 * h, b - 0.999 * h,
                   b + h, b - h]
    xs += [0.0, -0.0, 5e-324, 1e-300, 1e-8, 0.5, 1.0, -1.0, -3.0, math.pi, 2 * math.pi, 1e4, 99999.0]
    return xs


def test_dd_sincos_correctly_rounded(O):
    bad = []
    for x in _args():
        got = O.dd_sincos(x)
        want = exact_sincos(x)
        if got != want:
            bad.append((x, got, want))
    assert not bad, bad[:5]


def _exact_sin_cos_of(X):
    """sin and cos of the exact rational X (|X| <= pi/4) by the Taylor series."""
    def series(t, first):
        acc, n = F(0), first
        while t != 0 and (acc == 0 or abs(t) > abs(acc) / 2 ** 200):
            acc += t
            t = -t * X * X / ((n + 1) * (n + 2))
            n += 2
        return acc
    return series(X, 1), series(F(1), 0)


def test_dd_trig_table_entries(O):
    """sin and cos of j pi/128 in the table: the high parts correctly rounded, high +
    low within 2^-104 of the exact values (pi from Machin's formula)."""
    for j in range(33):
        sh, sl, ch, cl = O.dd_trig_table(j)
        S, Cv = _exact_sin_cos_of(j * PI / 128)
        assert sh == float(S) and ch == float(Cv), j
        for hi, lo, ex in ((sh, sl, S), (ch, cl, Cv)):
            if ex != 0:
                assert abs((F(hi) + F(lo) - ex) / ex) < F(1, 2 ** 104), j
            else:
                assert hi == 0.0 and lo == 0.0


def test_system_libm_is_a_last_bit_variant(O):
    """The C library's sin/cos, which the reference calls, agree with the correctly
    rounded values except in the last bit of a small fraction of arguments."""
    xs = _args()
    diff = 0
    for x in xs:
        want = exact_sincos(x)
        got = (math.sin(x), math.cos(x))
        if got != want:
            diff += 1
            for g, w in zip(got, want):
                if g != w:
                    assert abs(g - w) <= math.ulp(w)
    assert diff / len(xs) < 0.02
