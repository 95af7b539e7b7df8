"""Pins of the oracle's HEALPix primitives (vendored HEALPix 3.10 semantics:
pix_tools.F90, pixel_routines.F90).

Pinned by: the closed-form base-pixel centres of Gorski et al. (2005) at
nside 1, unit norm and ring structure of pix2vec_ring, ring_num inverting
ring2z, angdist against constructed angles, and query_disc against a brute
force scan of every pixel centre.
"""
import math

import numpy as np
import pytest

import p2mo


def all_vecs(nside):
    return np.array([p2mo.pix2vec_ring(nside, p) for p in range(12 * nside * nside)])


def test_nside2npix():
    L = p2mo.lib()
    for k in range(0, 14):
        assert L.p2mo_nside2npix(2 ** k) == 12 * 4 ** k
    for bad in (0, 3, 5, 6, 1000, -2):
        assert L.p2mo_nside2npix(bad) == -1


def test_nside1_base_pixels_closed_form():
    """Gorski et al. 2005: nside=1 centres at z=2/3 (phi=pi/4+k pi/2),
    z=0 (phi=k pi/2), z=-2/3 (phi=pi/4+k pi/2)."""
    v = all_vecs(1)
    exp = []
    for z, off in ((2 / 3, math.pi / 4), (0.0, 0.0), (-2 / 3, math.pi / 4)):
        s = math.sqrt(1 - z * z)
        for k in range(4):
            phi = off + k * math.pi / 2
            exp.append((s * math.cos(phi), s * math.sin(phi), z))
    assert np.allclose(v, np.array(exp), atol=1e-15)


@pytest.mark.parametrize("nside", [1, 2, 4, 16, 64])
def test_pix2vec_unit_and_rings(nside):
    v = all_vecs(nside)
    assert np.abs(np.linalg.norm(v, axis=1) - 1).max() < 2e-15
    # ring structure: ring i holds 4*min(i, nside, 4nside-i) pixels whose z is ring2z(i)
    p = 0
    for i in range(1, 4 * nside):
        npr = 4 * min(i, nside, 4 * nside - i)
        zr = p2mo.ring2z(nside, i)
        assert np.abs(v[p:p + npr, 2] - zr).max() < 1e-14
        phi = np.mod(np.arctan2(v[p:p + npr, 1], v[p:p + npr, 0]), 2 * np.pi)
        assert np.all(np.diff(phi) > 0)          # increasing azimuth along the ring
        assert np.allclose(np.diff(phi), 2 * np.pi / npr, atol=1e-12)
        p += npr
    assert p == 12 * nside * nside
    # equal-area pixelisation: centroid at the origin
    assert np.abs(v.mean(axis=0)).max() < 1e-12


@pytest.mark.parametrize("nside", [1, 4, 64, 1024])
def test_ring_num_inverts_ring2z(nside):
    for i in range(1, 4 * nside):
        assert p2mo.ring_num(nside, p2mo.ring2z(nside, i)) == i


def test_angdist_constructed():
    rng = np.random.default_rng(3)
    for _ in range(200):
        a = rng.normal(size=3)
        a /= np.linalg.norm(a)
        b = rng.normal(size=3)
        b -= a * (a @ b)
        b /= np.linalg.norm(b)
        alpha = rng.uniform(1e-6, math.pi - 1e-6)
        c = math.cos(alpha) * a + math.sin(alpha) * b
        d = p2mo.angdist(a, c)
        assert abs(d - alpha) < 1e-14 * max(1.0, alpha) + 1e-15
        assert p2mo.angdist(c, a) == pytest.approx(d, rel=1e-15)
        # scale invariance (the reference passes an un-normalised hpvec1, quirk Q2)
        assert p2mo.angdist(a / 3000.0, c) == pytest.approx(d, rel=1e-14)
    assert p2mo.angdist([0, 0, 1], [0, 0, 1]) == 0.0
    assert p2mo.angdist([0, 0, 1], [0, 0, -1]) == pytest.approx(math.pi, rel=1e-16)


def _brute(nside, v, r):
    V = all_vecs(nside)
    d = np.array([p2mo.angdist(v, w) for w in V])
    return d


@pytest.mark.parametrize("nside", [8, 16, 32])
def test_query_disc_equals_brute_force(nside):
    """query_disc returns exactly the pixel centres closer than the radius;
    disagreements are only possible within rounding of the edge."""
    rng = np.random.default_rng(nside)
    pix = 2 * math.sqrt(math.pi / 3) / nside / 2   # ~ pixel half-size (rad)
    centres = [np.array([0, 0, 1.0]), np.array([0, 0, -1.0]), np.array([1.0, 0, 0])]
    for _ in range(12):
        c = rng.normal(size=3)
        centres.append(c / np.linalg.norm(c))
    for c in centres:
        for r in (3 * pix, 10 * pix, 0.7, 2.0, math.pi):
            lst, gaps = p2mo.query_disc(nside, c, r)
            d = _brute(nside, c, r)
            inside = set(np.nonzero(d < r)[0].tolist())
            got = set(lst.tolist())
            assert len(got) == len(lst)                    # no duplicates
            for p in inside.symmetric_difference(got):
                assert abs(d[p] - r) < 1e-9, (p, d[p], r)
            assert gaps == 0


def test_query_disc_radius_checks():
    with pytest.raises(ValueError):
        p2mo.query_disc(8, [0, 0, 1.0], -0.1)
    with pytest.raises(ValueError):
        p2mo.query_disc(8, [0, 0, 1.0], 3.5)
    with pytest.raises(ValueError):
        p2mo.query_disc(8, [0, 0, 1.0], float("nan"))
    full, _ = p2mo.query_disc(8, [0.3, 0.2, 0.9], math.pi)
    assert sorted(full.tolist()) == list(range(12 * 64))
