"""Pins of the oracle's cosmology (cosmology.f90:51-214).

The distance tables are Euler sums, not exact integrals (quirk Q9), so they
are pinned against the exact flat-LCDM integral through the known leading
error of each scheme: rofz is a right-endpoint Riemann sum of dr/dz (error
+dz/2 (f(z)-f(0)) + O(dz^2), negative since f decreases); zofr is explicit (left-point) Euler on
dz/dr = H(z)/c, which underestimates z for increasing H.  A left/right
swap flips a sign and fails; a wrong step size fails the magnitude.
"""
import math

import numpy as np
import pytest
from scipy.integrate import quad

import p2mo

OM, OL, H = 0.25, 0.75, 0.7
CKMS = 2.99792458e10 / 1e5


@pytest.fixture(scope="module")
def o():
    hi = np.array([2, 2, 2], np.int32)
    hf = np.array([1e2, 7e3, 1e13, 1e16, 1e-2, 10], np.float32)
    return p2mo.Oracle(hi, hf, np.zeros(16, np.float32))


def f_drdz(z):
    return CKMS / (100 * H) / math.sqrt(OM * (1 + z) ** 3 + 1 - OM)


def chi_exact(z):
    return quad(f_drdz, 0, z, epsabs=0, epsrel=1e-13)[0]


def test_rofz_right_riemann_sum(o):
    dz = 6.0 / 9999
    for z in (0.05, 0.3, 0.7, 1.2, 2.5):
        # evaluate at table nodes to isolate the quadrature error
        i = round(z / dz)
        zn = i * dz
        expected = chi_exact(zn) + 0.5 * dz * (f_drdz(zn) - f_drdz(0.0))
        assert o.rofz(zn) == pytest.approx(expected, rel=2e-7)
        assert o.rofz(zn) < chi_exact(zn)


def test_zofr_forward_euler(o):
    # z(chi) by inverting the exact integral
    for chi in (200.0, 1000.0, 3000.0, 3900.0):
        zt = 0.0
        lo, hi = 0.0, 6.0
        for _ in range(80):
            mid = 0.5 * (lo + hi)
            if chi_exact(mid) < chi:
                lo = mid
            else:
                hi = mid
        zt = 0.5 * (lo + hi)
        got = o.zofr(chi)
        assert got < zt                     # explicit Euler underestimates
        assert abs(got - zt) < 2e-3 * zt + 1e-6


def test_round_trip_and_monotone(o):
    z = np.linspace(0.01, 5.0, 500)
    chi = np.array([o.rofz(x) for x in z])
    assert np.all(np.diff(chi) > 0)
    back = np.array([o.zofr(c) for c in chi])
    assert np.abs(back - z).max() < 5e-3


def test_table_end_clamps(o):
    assert o.zofr(-5.0) == 0.0
    assert o.rofz(-1.0) == 0.0
    assert o.rofz(7.0) == o.rofz(6.5)
    assert o.zofr(2.0e4) == o.zofr(1.9e4)


def test_rhocrit_and_deltacrit(o):
    rc0 = 2.775e11 * H * H
    assert o.rhocrit(0.0) == pytest.approx(rc0, rel=1e-15)
    assert o.rhocrit(1e4) == pytest.approx(rc0 * OM, rel=1e-6)
    # Bryan & Norman (1998): 18 pi^2 + 82 x - 39 x^2 with x = Omega_m(z) - 1,
    # pi being the float-rounded constant of cosmology.f90:8-9 (quirk Q8)
    pif = float(np.float32(3.14159265359))
    assert o.deltacrit(1e4) == pytest.approx(18 * pif ** 2, rel=1e-9)
    x = OM - 1
    assert o.deltacrit(0.0) == pytest.approx(18 * pif ** 2 + 82 * x - 39 * x * x, rel=1e-14)
    assert 90 < o.deltacrit(0.0) < 100
