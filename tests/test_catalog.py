"""The synthetic catalogue generator: deterministic, chunk-independent, and
drawn from the stated distributions."""
import numpy as np
import pytest
from scipy.integrate import quad

from matt_young_mapmaking_amd import catalog

C_KMS = 299792.458


def test_deterministic_and_prefix_consistent():
    a = catalog.make_catalog(5000, seed=13579)
    b = catalog.make_catalog(5000, seed=13579)
    assert np.array_equal(a, b)
    big = catalog.make_catalog(3_000_000, seed=13579)
    part = catalog.make_catalog(100, seed=13579, start=(1 << 20) - 50)
    assert np.array_equal(part, big[(1 << 20) - 50:(1 << 20) + 50])
    assert not np.array_equal(a, catalog.make_catalog(5000, seed=1))


def test_comoving_distance_matches_quadrature():
    for z in (0.05, 0.5, 1.2):
        exact = C_KMS / 68.0 * quad(lambda x: 1 / np.sqrt(0.31 * (1 + x) ** 3 + 0.69), 0, z)[0]
        assert catalog.comoving_distance(z) == pytest.approx(exact, rel=1e-6)
    assert catalog.comoving_distance(0.05) == pytest.approx(217.8, abs=0.1)
    assert catalog.comoving_distance(1.2) == pytest.approx(3848.7, abs=0.5)


def test_distributions():
    h = catalog.make_catalog(400_000, geom="sky", mlo=1e13, mhi=1e15).astype(np.float64)
    chi = np.linalg.norm(h[:, :3], axis=1)
    lo, hi = catalog.comoving_distance([0.05, 1.2])
    assert chi.min() >= lo * (1 - 1e-6) and chi.max() <= hi * (1 + 1e-6)
    # uniform in volume: chi^3 uniform
    u = (chi ** 3 - lo ** 3) / (hi ** 3 - lo ** 3)
    assert abs(u.mean() - 0.5) < 0.005
    # isotropic
    assert np.abs((h[:, :3] / chi[:, None]).mean(axis=0)).max() < 0.005
    # masses from Rth
    rhobar = 2.775e11 * 0.31 * 0.68 ** 2
    m = 4 / 3 * np.pi * rhobar * h[:, 6] ** 3
    assert m.min() >= 1e13 * (1 - 1e-5) and m.max() <= 1e15 * (1 + 1e-5)
    # dn/dM ~ M^-1.9: median of the truncated power law
    g = -0.9
    med = (0.5 * (1e13 ** g + 1e15 ** g)) ** (1 / g)
    assert np.median(m) == pytest.approx(med, rel=0.02)
    assert abs(np.std(h[:, 3:6]) - 300) < 3


def test_flat_patch_geometry():
    h = catalog.make_catalog(100_000, geom="patch").astype(np.float64)
    chi = np.linalg.norm(h[:, :3], axis=1)
    th = np.degrees(np.arcsin(h[:, 2] / chi))     # painter x = pos3
    ph = np.degrees(np.arcsin(h[:, 1] / chi))
    assert np.abs(th).max() <= 5.0 + 1e-4 and np.abs(ph).max() <= 5.0 + 1e-4
    assert (h[:, 0] > 0).all()                     # painter z = pos1 > 0
