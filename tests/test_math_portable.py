"""Pins of the portable FP64 elementary functions (p2m_math.h) that both the
HIP kernels and the oracle use.

What pins them: agreement with glibc's libm to a few ulp over the domains the
painter uses, exact special values fixed by the mathematics, and identities
(sin^2+cos^2=1, exp(log x)=x, atan2 quadrant/sign rules).  A wrong Taylor
coefficient, a wrong Cody-Waite constant or a wrong quadrant fails these.
"""
import math

import numpy as np
import pytest

import p2mo


def ulp_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    sp = np.spacing(np.abs(b))
    sp[sp == 0] = np.finfo(np.float64).tiny
    return np.abs(a - b) / sp


RNG = np.random.default_rng(7)


@pytest.mark.parametrize("op,gen,ref,maxulp", [
    ("sin", lambda n: RNG.uniform(-7, 7, n), np.sin, 2),
    ("cos", lambda n: RNG.uniform(-7, 7, n), np.cos, 2),
    ("tan", lambda n: RNG.uniform(-1.5, 1.5, n), np.tan, 3),
    ("asin", lambda n: RNG.uniform(-1, 1, n), np.arcsin, 3),
    ("acos", lambda n: RNG.uniform(-1, 1, n), np.arccos, 3),
    ("log", lambda n: np.exp(RNG.uniform(-40, 40, n)), np.log, 2),
    ("exp", lambda n: RNG.uniform(-700, 700, n), np.exp, 2),
])
def test_unary_vs_libm(op, gen, ref, maxulp):
    x = gen(200000)
    got = p2mo.math(op, x)
    assert ulp_err(got, ref(x)).max() <= maxulp


def test_atan2_vs_libm():
    y = RNG.uniform(-3, 3, 200000)
    x = RNG.uniform(-3, 3, 200000)
    assert ulp_err(p2mo.math("atan2", y, x), np.arctan2(y, x)).max() <= 3


def test_pow_cuberoot_vs_libm():
    x = np.exp(RNG.uniform(-30, 30, 200000))
    got = p2mo.math("pow", x, np.full_like(x, 1.0 / 3.0))
    assert ulp_err(got, np.power(x, 1.0 / 3.0)).max() <= 8


def test_sqrt_is_correctly_rounded():
    x = np.concatenate([np.exp(RNG.uniform(-50, 50, 100000)),
                        np.arange(1, 5000, dtype=np.float64) ** 2,
                        np.nextafter(np.arange(1, 5000, dtype=np.float64) ** 2, 0)])
    assert np.array_equal(p2mo.math("sqrt", x), np.sqrt(x))


def test_special_values():
    assert p2mo.math("sin", [0.0])[0] == 0.0
    assert p2mo.math("cos", [0.0])[0] == 1.0
    assert p2mo.math("log", [1.0])[0] == 0.0
    assert p2mo.math("exp", [0.0])[0] == 1.0
    assert p2mo.math("asin", [1.0])[0] == math.pi / 2
    assert p2mo.math("asin", [-1.0])[0] == -math.pi / 2
    assert p2mo.math("acos", [-1.0])[0] == math.pi
    assert p2mo.math("acos", [1.0])[0] == 0.0
    # atan2 quadrants and signed zeros (IEEE conventions)
    y = np.array([0.0, 0.0, -0.0, -0.0, 1.0, -1.0, 1.0, -1.0])
    x = np.array([1.0, -1.0, 1.0, -1.0, 0.0, 0.0, -1.0, -1.0])
    got = p2mo.math("atan2", y, x)
    exp = np.arctan2(y, x)
    assert np.array_equal(np.signbit(got), np.signbit(exp))
    assert np.allclose(got, exp, rtol=0, atol=1e-16)
    assert np.isnan(p2mo.math("asin", [1.5])[0])
    assert np.isneginf(p2mo.math("log", [0.0])[0])


def test_identities():
    x = RNG.uniform(-10, 10, 100000)
    s = p2mo.math("sin", x)
    c = p2mo.math("cos", x)
    assert np.abs(s * s + c * c - 1).max() < 5e-16
    t = p2mo.math("tan", x)
    ok = np.abs(c) > 1e-3
    assert np.allclose(t[ok], s[ok] / c[ok], rtol=1e-14)
    u = np.exp(RNG.uniform(-20, 20, 100000))
    assert np.allclose(p2mo.math("exp", p2mo.math("log", u)), u, rtol=1e-14)


def test_portable_flavour_differs_from_libm_flavour():
    """The two oracle builds really use different implementations (otherwise
    the portable-vs-libm membership comparison would prove nothing)."""
    x = RNG.uniform(-7, 7, 100000)
    assert not np.array_equal(p2mo.math("sin", x, flavour="portable"), p2mo.math("sin", x, flavour="libm"))
