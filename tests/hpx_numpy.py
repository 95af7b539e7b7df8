"""Vectorised HEALPix RING pixel centres from the closed forms of Gorski et
al. (2005, ApJ 622, 759, eqs. 4-9): used by the tests to preselect candidate
pixels for brute-force checks (final decisions use the oracle's exact
primitives)."""
import numpy as np


def ring_of_pixels(nside):
    """(ring index 1..4n-1, in-ring index, npr) for every RING pixel."""
    rings = np.arange(1, 4 * nside)
    npr = 4 * np.minimum(np.minimum(rings, nside), 4 * nside - rings)
    ring = np.repeat(rings, npr)
    start = np.concatenate([[0], np.cumsum(npr)[:-1]])
    ip = np.arange(12 * nside * nside) - np.repeat(start, npr)
    return ring, ip, np.repeat(npr, npr)


def pix2vec(nside):
    ring, ip, npr = ring_of_pixels(nside)
    n = float(nside)
    z = np.where(ring < nside, 1 - ring ** 2 / (3 * n * n),
                 np.where(ring <= 3 * nside, (2 * n - ring) * 2 / (3 * n),
                          -1 + (4 * n - ring) ** 2 / (3 * n * n)))
    shift = np.where((ring < nside) | (ring > 3 * nside), 0.5, 0.5 * ((ring + nside + 1) % 2))
    phi = (ip + shift) * 2 * np.pi / npr
    s = np.sqrt(np.maximum(0.0, 1 - z * z))
    return np.stack([s * np.cos(phi), s * np.sin(phi), z], axis=1)
