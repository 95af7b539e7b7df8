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


# physical constants of cosmology.f90:11-23 (double parameters; mue and mp are
# default-real literals, float-rounded) and the pks2map/make_maptable cosmology
SIGMAT, C_CGS, ME, MSUN = 6.65246e-25, 2.99792458e10, 9.109383e-28, 1.9891e33
MPC2KM, MPC, KPC = 3.08567758e19, 3.08567758e24, 3.08567758e21
F32 = lambda v: float(np.float32(v))
PI_F = F32(3.14159265359)
OM, OL, H, OB = 0.25, 0.75, 0.7, 0.043


def node(o, ic, im, ir):
    """(z, mh, b, chit) of table node (ir, im, ic) of the NCHI x NM x NR grid"""
    lc0, dch = math.log(F32(1e2)), (math.log(F32(7e3)) - math.log(F32(1e2))) / (NCHI - 1)
    lm0, dm = math.log(F32(1e13)), (math.log(F32(1e16)) - math.log(F32(1e13))) / (NM - 1)
    lr0, dr = math.log(F32(1e-2)), (math.log(10.0) - math.log(F32(1e-2))) / (NR - 1)
    chit = math.exp(lc0 + (ic - 1) * dch)
    z = o.zofr(chit)
    mh = math.exp(lm0 + (im - 1) * dm)
    rvir = (3 * mh / 4 / PI_F / 200 / o.rhocrit(z)) ** (1 / 3)
    rt = math.exp(lr0 + (ir - 1) * dr)
    b = o.rofz(z) * math.sin(math.atan(rt / chit)) / rvir
    return z, mh, b, chit


def test_y_amplitude_across_nodes(tabs):
    """The y table over the pressure integral is y0 M E^2(z) at every node,
    with y0 = fb Delta h^2 * 3 sigma_T H100^2 / (8 pi me c^2) Msun / 1.932
    (maptable.f90:60, cosmology.f90:25-28) and E^2 = Om (1+z)^3 + OL: the
    mass and redshift scaling across (chi, M) nodes, not only within one."""
    o = p2mo.Oracle(*tabs[0.0])
    t = tabs[0.0][2]
    h100 = 1e2 / MPC2KM
    ypermass = 3 * SIGMAT * h100 ** 2 / 8.0 / PI_F / ME / C_CGS ** 2 * MSUN / F32(1.932)
    y0 = (F32(OB) / F32(OM)) * 200.0 * F32(H) ** 2 * ypermass
    seen = 0
    for ic in range(1, NCHI + 1):
        for im in range(1, NM + 1):
            for ir in (8, 16, 24, 32):
                z, mh, b, chit = node(o, ic, im, ir)
                if not (0.2 < b < 3.5):
                    continue
                P0 = F32(18.1) * (mh / 1e14) ** F32(0.154) * (1 + z) ** F32(-0.758)
                xc = F32(0.497) * (mh / 1e14) ** F32(-0.00865) * (1 + z) ** F32(0.731)
                beta = F32(4.35) * (mh / 1e14) ** F32(0.0393) * (1 + z) ** F32(0.415)
                P = lambda x: P0 * (x / xc) ** F32(-0.3) * (1 + x / xc) ** (-beta)
                smax = (int(4 / F32(1e-2)) + 1) * F32(1e-2)
                I = quad(lambda s: P(math.sqrt(s * s + b * b)), 0, smax, epsrel=1e-12, limit=200)[0]
                want = y0 * mh * (F32(OM) * (1 + z) ** 3 + F32(OL)) * I
                assert float(T(t, 1, ir, im, ic)) == pytest.approx(want, rel=5e-5), (ic, im, ir)
                seen += 1
    assert seen >= 40


def test_tau_profile_against_quadrature(tabs):
    """tau = tau0 M^(1/3) zfact * 2 * the line-of-sight integral of the
    published gas density fit (bbps_profile.f90:140-180) cut at the sphere of
    4 rvir (s0 = sqrt(16 - b^2), n = int(s0/ds)+1 midpoint cells, so the sum
    spans [0, n ds]), with tau0 = sigma_T / mue / mp Msun / kpc^3 Mpc 1e10 h^2
    (3 / (4 pi Delta))^(1/3) and zfact = 1 / (1+z) / rho_c(z)^(1/3)
    (maptable.f90:61, 77-79; integrate_profiles.f90:56-95).  A cylindrical cut
    or a missing factor 2 misses by far more than the midpoint rule's error."""
    o = p2mo.Oracle(*tabs[0.0])
    t = tabs[0.0][2]
    taupersigma = SIGMAT / F32(1.136) / F32(1.67e-24) * MSUN / KPC ** 3 * MPC * 1e10
    tau0 = taupersigma * F32(H) ** 2 * (3.0 / 4.0 / PI_F / 200.0) ** (1.0 / 3.0)
    ds = F32(1e-2)
    seen = 0
    for ic in range(1, NCHI + 1):
        for im in range(1, NM + 1):
            for ir in (8, 16, 24, 32):
                z, mh, b, chit = node(o, ic, im, ir)
                if not (0.2 < b < 3.5):
                    continue
                q = (mh / 1e14, 1 + z)
                P0 = F32(1.9e-5) * q[0] ** F32(0.29) * q[1] ** F32(-0.67)
                al = F32(0.88) * q[0] ** F32(-0.03) * q[1] ** F32(0.20)
                be = F32(3.83) * q[0] ** F32(0.044) * q[1] ** F32(-0.025)
                g = F32(0.2)
                rho = lambda x: P0 * (x / 0.5) ** (-g) * (1 + (x / 0.5) ** al) ** (-(be - g) / al)
                s0 = math.sqrt(16.0 - b * b)
                send = (int(s0 / ds) + 1) * ds
                I = 2 * quad(lambda s: rho(math.sqrt(s * s + b * b)), 0, send, epsrel=1e-12, limit=200)[0]
                zfact = 1.0 / (1 + z) / o.rhocrit(z) ** (1.0 / 3.0)
                want = tau0 * mh ** (1.0 / 3.0) * zfact * I
                assert float(T(t, 2, ir, im, ic)) == pytest.approx(want, rel=5e-5), (ic, im, ir)
                seen += 1
    assert seen >= 40
