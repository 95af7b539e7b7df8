"""Pins of the oracle's make_maptable (maptable.f90:25-146,
integrate_profiles.f90:12-95, bbps_profile.f90:60-180).

  * frequency: y(nu)/y(nu') at one node equals G(x)/G(x') with
    G(x) = x coth(x/2) - 4, x = 0.017612 nu (1+z)^alpha, computed here from the
    definition; tau does not depend on nu; G changes sign near 217.5 GHz;
  * line-of-sight integral: the midpoint sum of the Battaglia et al. (2012)
    pressure fit with n = int(4/ds)+1 = 401 cells of ds = 1e-2 (so it spans
    s in [0, 4.01], cylindrical cut) matches an adaptive quadrature of the
    published fit over that span to the midpoint rule's accuracy, node by
    node in r (ratios, so the normalisation constants cancel).  A 400-cell
    sum or an end-point rule fails.
"""
import math

import numpy as np
import pytest
from scipy.integrate import quad

import p2mo

NCHI, NM, NR = 6, 5, 40


def G(x):
    return x * (math.exp(x) + 1) / (math.exp(x) - 1) - 4


@pytest.fixture(scope="module")
def tabs():
    out = {}
    for nu in (0.0, 148.0, 218.0, 353.0):
        out[nu] = p2mo.make_maptable(nu=nu, alpha=0.0, nchit=NCHI, nmht=NM, nrt=NR, nthreads=8)
    return out


def T(t, p, ir, im, ic):
    return t[(p - 1) + 2 * ((ir - 1) + NR * ((im - 1) + NM * (ic - 1)))]


def test_header(tabs):
    hi, hf, t = tabs[148.0]
    assert list(hi) == [NCHI, NM, NR]
    assert np.array_equal(hf, np.array([1e2, 7e3, 1e13, 1e16, 1e-2, 10], np.float32))
    assert t.size == 2 * NCHI * NM * NR


def test_frequency_factor(tabs):
    t0 = tabs[0.0][2]
    for nu in (148.0, 218.0, 353.0):
        t = tabs[nu][2]
        x = 0.017612 * nu
        for (ir, im, ic) in ((3, 2, 2), (20, 4, 5), (35, 1, 3)):
            y0 = float(T(t0, 1, ir, im, ic))
            if y0 == 0.0:
                continue
            assert float(T(t, 1, ir, im, ic)) == pytest.approx(y0 * G(x), rel=2e-6)
        assert np.array_equal(t[1::2], t0[1::2])         # tau is frequency independent
    assert G(0.017612 * 217.0) < 0 < G(0.017612 * 218.0)
    assert (tabs[148.0][2][0::2] <= 0).all() and (tabs[353.0][2][0::2] >= 0).all()


def test_midpoint_los_integral_against_quadrature(tabs):
    o = p2mo.Oracle(*tabs[0.0])
    t = tabs[0.0][2]
    lc0, dch = math.log(float(np.float32(1e2))), (math.log(float(np.float32(7e3))) - math.log(float(np.float32(1e2)))) / (NCHI - 1)
    lm0, dm = math.log(float(np.float32(1e13))), (math.log(float(np.float32(1e16))) - math.log(float(np.float32(1e13)))) / (NM - 1)
    lr0, dr = math.log(float(np.float32(1e-2))), (math.log(10.0) - math.log(float(np.float32(1e-2)))) / (NR - 1)
    pif = float(np.float32(3.14159265359))
    for ic, im in ((3, 2), (5, 4), (2, 3)):
        chit = math.exp(lc0 + (ic - 1) * dch)
        z = o.zofr(chit)
        mh = math.exp(lm0 + (im - 1) * dm)
        chi = o.rofz(z)
        rvir = (3 * mh / 4 / pif / 200 / o.rhocrit(z)) ** (1 / 3)
        # Battaglia et al. 2012 (AGN feedback, Delta=200) fit, float-rounded as in bbps_profile.f90:9-19
        f = lambda v: float(np.float32(v))
        P0 = f(18.1) * (mh / 1e14) ** f(0.154) * (1 + z) ** f(-0.758)
        xc = f(0.497) * (mh / 1e14) ** f(-0.00865) * (1 + z) ** f(0.731)
        beta = f(4.35) * (mh / 1e14) ** f(0.0393) * (1 + z) ** f(0.415)
        gam = f(-0.3)
        P = lambda x: P0 * (x / xc) ** gam * (1 + x / xc) ** (-beta)
        ratios = []
        for ir in range(1, NR + 1):
            rt = math.exp(lr0 + (ir - 1) * dr)
            b = chi * math.sin(math.atan(rt / chit)) / rvir
            if not (0.2 < b < 3.5):
                continue
            ds = float(np.float32(1e-2))
            smax = (int(4 / ds) + 1) * ds
            I = quad(lambda s: P(math.sqrt(s * s + b * b)), 0, smax, epsrel=1e-12, limit=200)[0]
            ratios.append(float(T(t, 1, ir, im, ic)) / I)
        ratios = np.array(ratios)
        assert len(ratios) >= 5
        # y = y0 * M * E^2(z) * integral: one constant per (chi, M) node
        assert ratios.std() / abs(ratios.mean()) < 2e-5
