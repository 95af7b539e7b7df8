"""The oracle's `ranf` (random.f90:40-57 with ranfmodmult :264-286) against
the closed form of the generator it implements.

Pinned by: the four 12-bit limbs (s1 least significant) are one 48-bit
integer X = s1 + 4096 s2 + 4096^2 s3 + 4096^3 s4; a draw returns X / 2^48 and
the next state is (A X) mod 2^48 with A = 373 + 3707*4096 + 1442*4096^2 +
647*4096^3 (the Multiplier limbs, random.f90:28-29).  Python integers give
that exactly, so a dropped carry, a swapped limb or a wrong divisor fails.
The scramble seed of pks2map.f90:164 sets every limb to 13579 > 4096
(quirk Q10): the first draw is 13579 * (1/4096 + ... + 1/4096^4) > 1, and the
carry arithmetic then reduces the state to valid limbs.
"""
import numpy as np
import pytest

import p2mo

A = 373 + 3707 * 4096 + 1442 * 4096 ** 2 + 647 * 4096 ** 3
M48 = 1 << 48


def draw(s):
    """one draw through the oracle; updates the limb list in place"""
    r, new = p2mo.ranf(s)
    s[:] = new
    return r


def limbs_to_int(s):
    return s[0] + 4096 * s[1] + 4096 ** 2 * s[2] + 4096 ** 3 * s[3]


@pytest.mark.parametrize("seed", [(3281, 4041, 595, 2376), (1, 0, 0, 0), (4095, 4095, 4095, 4095), (17, 2, 3000, 9)])
def test_ranf_is_the_48bit_lcg(seed):
    s = list(seed)
    x = limbs_to_int(s)
    for _ in range(2000):
        r = draw(s)
        assert r == x / M48            # exact: every term is a dyadic fraction
        x = (A * x) % M48
        assert limbs_to_int(s) == x and all(0 <= v < 4096 for v in s)


def test_scramble_seed_quirk():
    s = [13579] * 4
    r0 = draw(s)
    expect = 13579 * (1 / 4096 + 1 / 4096 ** 2 + 1 / 4096 ** 3 + 1 / 4096 ** 4)
    assert r0 == pytest.approx(expect, rel=1e-15) and r0 > 3.3
    x = (A * (13579 * (1 + 4096 + 4096 ** 2 + 4096 ** 3))) % M48
    assert limbs_to_int(s) == x
    r1 = draw(s)
    assert r1 == x / M48 and 0 <= r1 < 1


def test_ranf_uniformity():
    s = [3281, 4041, 595, 2376]
    r = np.array([draw(s) for _ in range(20000)])
    assert 0 <= r.min() and r.max() < 1
    hist = np.histogram(r, bins=10, range=(0, 1))[0]
    assert np.all(np.abs(hist - 2000) < 5 * np.sqrt(2000))
