"""The oracle's host branches of pks2map (radial cut, scramble, centring;
pks2map.f90:144-262) pinned by what the geometry fixes:

  * the cut keeps exactly the halos whose float distance lies in the band, in
    catalogue order, records intact;
  * scramble keeps every halo's distance, draws mu then phi from the 48-bit
    LCG in that order (mu of halo i is 2 ranf_{2i} - 1 of the closed-form
    generator), and the Q10 seed makes the first halo NaN;
  * centring is sign flips, a rotation and an x-z exchange: norms and the
    flipped catalogue's pairwise dot products are preserved, and the largest halo ends on the +pos(1)
    axis, the painter's z axis (pks2map.f90:240-244).
"""
import numpy as np
import pytest

import p2mo

A = 373 + 3707 * 4096 + 1442 * 4096 ** 2 + 647 * 4096 ** 3
M48 = 1 << 48


def catalogue(n, seed=5):
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    chi = rng.uniform(100.0, 3000.0, n)
    h = np.zeros((n, 7), np.float32)
    h[:, 0:3] = d * chi[:, None]
    h[:, 3:6] = rng.normal(0, 300, (n, 3))
    h[:, 6] = rng.uniform(0.5, 3.0, n)
    return h


def test_cut_keeps_band_in_order():
    h = catalogue(5000)
    lo, hi, chv = np.float32(800.0), np.float32(2200.0), np.float32(50.0)
    out = p2mo.host_branches(h, lo, hi, chv)
    x = h[:, 2] - chv
    chi = np.sqrt((x * x + h[:, 1] * h[:, 1]) + h[:, 0] * h[:, 0])
    keep = ~((chi > hi) | (chi < lo))
    assert np.array_equal(out, h[keep])
    assert 0 < keep.sum() < h.shape[0]


def test_scramble_keeps_distance_and_follows_the_generator():
    h = catalogue(3000)
    out = p2mo.host_branches(h, 0.0, 1e9, 0.0, scramble=1)
    assert out.shape == h.shape
    assert np.all(np.isnan(out[0, :2])) and out[0, 2] > 5 * np.linalg.norm(h[0, :3])  # Q10: mu > 1
    r0 = np.linalg.norm(h[1:, :3].astype(np.float64), axis=1)
    r1 = np.linalg.norm(out[1:, :3].astype(np.float64), axis=1)
    assert np.max(np.abs(r1 / r0 - 1)) < 1e-6
    assert np.array_equal(out[:, 3:], h[:, 3:])
    # closed-form LCG: draws 2i (mu) and 2i+1 (phi) of the stream seeded with 13579 limbs
    x = 13579 * (1 + 4096 + 4096 ** 2 + 4096 ** 3)
    draws = []
    for k in range(2 * h.shape[0]):
        draws.append(x / M48 if k else 13579 * (1 / 4096 + 1 / 4096 ** 2 + 1 / 4096 ** 3 + 1 / 4096 ** 4))
        x = (A * x) % M48
    draws = np.array(draws)
    mu = (2 * draws[0::2] - 1).astype(np.float32)
    phi = (np.float64(np.float32(2) * np.float32(3.14159)) * draws[1::2]).astype(np.float32)
    chi = np.sqrt((h[1:, 0] * h[1:, 0] + h[1:, 1] * h[1:, 1]) + h[1:, 2] * h[1:, 2])
    assert np.array_equal(out[1:, 2], mu[1:] * chi)
    ang = np.arctan2(out[1:, 1].astype(np.float64), out[1:, 0].astype(np.float64)) % (2 * np.pi)
    dphi = np.abs((ang - phi[1:].astype(np.float64) + np.pi) % (2 * np.pi) - np.pi)
    ok = np.abs(mu[1:]) < 0.999
    assert np.max(dphi[ok]) < 1e-4
    # isotropy of the kept draws
    assert abs(np.mean(mu[1:])) < 0.05 and abs(np.mean(mu[1:] ** 2) - 1 / 3) < 0.03


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_centre_puts_largest_halo_on_axis(seed):
    h = catalogue(2000, seed)
    out = p2mo.host_branches(h, 0.0, 1e9, 0.0, centre=1)
    k = int(np.argmax(h[:, 6]))
    r = np.linalg.norm(h[k, :3].astype(np.float64))
    assert out[k, 0] > 0 and abs(out[k, 0] / r - 1) < 1e-5
    assert abs(out[k, 1]) < 1e-5 * r and abs(out[k, 2]) < 1e-5 * r
    sgn = np.where(h[k, :3] < 0, -1.0, 1.0)
    a = h[:, :3].astype(np.float64) * sgn
    b = out[:, :3].astype(np.float64)
    assert np.max(np.abs(np.linalg.norm(b, axis=1) / np.linalg.norm(a, axis=1) - 1)) < 1e-5
    i = np.arange(0, 2000, 7)
    ga, gb = a[i] @ a[i].T, b[i] @ b[i].T
    assert np.max(np.abs(gb - ga)) < 1e-5 * np.max(np.abs(ga))
    # a rotation in painter axes, then stored with x and z exchanged (pks2map.f90:240-242):
    # one reflection in all
    R, *_ = np.linalg.lstsq(a, b, rcond=None)
    assert np.linalg.det(R) == pytest.approx(-1.0, abs=1e-4)


def test_scramble_then_centre_and_empty():
    h = catalogue(500)
    out = p2mo.host_branches(h, 0.0, 1e9, 0.0, scramble=1, centre=1)
    assert out.shape == h.shape
    assert p2mo.host_branches(h, 1e8, 1e9, 0.0, scramble=1, centre=1).shape == (0, 7)
