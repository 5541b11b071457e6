"""The FP64 demodulation sin/cos of the column pass (sincos_cw in pgw_kernels.hpp): its reduction
constants against the values the mathematics fixes (pi/2 split in three doubles, 2/pi); its
polynomials, evaluated in 200-bit arithmetic on their double coefficients, against sin and cos on
the reduced range (the minimax error bound, which a slip in a leading digit of any coefficient
or a dropped term exceeds); and the whole reduction and evaluation, with exact FMAs on the same
constants, against 200-bit sin and cos.  The device code itself is exercised by the FP64 parity
tests; this pins what a plausible slip would change by far less than their 1e-9 tolerance on F."""
import math
import random
import re
from fractions import Fraction as Fr
from pathlib import Path

import pytest

mpmath = pytest.importorskip("mpmath")

SRC = Path(__file__).resolve().parents[1] / "polgraw-allsky_amd" / "csrc" / "pgw_kernels.hpp"


def _body():
    s = SRC.read_text()
    i = s.index("void sincos_cw(")
    return s[i:s.index("\n}\n", i)]


def _constants():
    b = _body()
    m = re.search(r"P1 = ([-0-9.e]+), P2 = ([-0-9.e]+), P3 = ([-0-9.e]+);", b)
    p = [float(x) for x in m.groups()]
    two_pi = float(re.search(r"rint\(x \* ([0-9.e]+)\)", b).group(1))
    ps = [float(x) for x in re.findall(r"ps = (?:fma\(ps, r2, )?([-0-9.e]+)\)?;", b)]
    pc = [float(x) for x in re.findall(r"pc = (?:fma\(pc, r2, )?([-0-9.e]+)\)?;", b)]
    return p, two_pi, ps, pc


def test_constants_are_the_mathematical_ones():
    mpmath.mp.prec = 300
    (p1, p2, p3), two_pi, ps, pc = _constants()
    half_pi = mpmath.pi / 2
    # the split: each term the double nearest to what the previous ones leave of pi/2
    assert p1 == float(half_pi)
    assert p2 == float(half_pi - p1)
    assert p3 == float(half_pi - p1 - p2)
    assert abs(half_pi - p1 - p2 - p3) < mpmath.mpf(2) ** -160
    assert two_pi == float(2 / mpmath.pi)


def test_polynomials_on_the_reduced_range():
    """sin r = r + r^3 S(r^2), cos r = 1 - r^2 C(r^2) on |r| <= pi/4, in exact arithmetic on the
    kernel's double coefficients: within the minimax bounds (1.6e-18, 5.2e-17) plus the
    coefficients' rounding.  A coefficient off in its 12th significant digit, or a term dropped,
    moves the error far past them; S and C start at -1/3! and 1/2! (to 1e-13), as the series
    fix."""
    mpmath.mp.prec = 200
    _, _, ps, pc = _constants()
    assert len(ps) == 6 and len(pc) == 6
    assert abs(ps[-1] + 1 / 6) < 1e-13 and abs(pc[-1] - 0.5) < 1e-13
    es = ec = mpmath.mpf(0)
    n = 2000
    for i in range(n + 1):
        r = mpmath.pi / 4 * i / n
        z = r * r
        S = mpmath.mpf(0)
        for c in ps:
            S = S * z + c
        C = mpmath.mpf(0)
        for c in pc:
            C = C * z + c
        es = max(es, abs(r + r * z * S - mpmath.sin(r)))
        ec = max(ec, abs(1 - z * C - mpmath.cos(r)))
    assert es < 1e-17, es
    assert ec < 7e-17, ec


def _fma(a, b, c):
    return float(Fr(a) * Fr(b) + Fr(c))


def _sincos(x, consts):
    (p1, p2, p3), two_pi, ps, pc = consts
    k = float(round(x * two_pi))  # rint: ties to even, as round()
    r = _fma(-k, p1, x)
    r = _fma(-k, p2, r)
    r = _fma(-k, p3, r)
    r2 = r * r
    s = ps[0]
    for c in ps[1:]:
        s = _fma(s, r2, c)
    sn = _fma(r * r2, s, r)
    c = pc[0]
    for d in pc[1:]:
        c = _fma(c, r2, d)
    cs = _fma(-r2, c, 1.0)
    q = int(_fma(-4.0, math.floor(k * 0.25), k))
    s0, c0 = (cs, sn) if q & 1 else (sn, cs)
    return (-s0 if q & 2 else s0), (-c0 if (q + 1) & 2 else c0)


def test_reduction_and_series_accuracy():
    mpmath.mp.prec = 200
    consts = _constants()
    rng = random.Random(7)
    xs = [0.0, 1e-300, math.pi / 4, math.pi / 2, -math.pi, 3 * math.pi / 4, 1e7, -1e7, 2.0 ** 40]
    xs += [rng.uniform(-sc, sc) for sc in (4, 1e3, 1e5, 1e7, 1e9, 1e12) for _ in range(150)]
    xs += [k * math.pi / 4 + d for k in range(-8, 9) for d in (-1e-9, 0.0, 1e-9)]  # octant edges
    worst = 0.0
    for x in xs:
        s, c = _sincos(x, consts)
        es = abs(s - float(mpmath.sin(mpmath.mpf(x))))
        ec = abs(c - float(mpmath.cos(mpmath.mpf(x))))
        worst = max(worst, es, ec)
    # ~1 ulp of values below 1 (the demodulation phases here stay below 1e7)
    assert worst <= 2.5e-16, worst
