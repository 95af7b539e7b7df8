"""File layouts of every file the path reads or writes (formats.py), pinned by
hand-built byte fixtures (tests/golden/make_golden.py) and round trips."""
import os
import struct

import numpy as np
import pytest

from matt_young_mapmaking_amd import formats

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_pksc_golden_fixture():
    n, rthmax, h = formats.read_pksc(os.path.join(GOLD, "three_halos.pksc"))
    assert n == 3
    assert rthmax == 12.5
    exp = np.array([
        (1200.5, -33.25, 450.0, 120.0, -75.5, 10.0, 3.5),
        (-2500.0, 1800.75, -90.125, -300.0, 260.0, 5.25, 7.0),
        (0.0, 0.0, 3000.0, 0.0, 0.0, 0.0, 12.5)], np.float32)
    assert np.array_equal(h, exp)
    raw = open(os.path.join(GOLD, "three_halos.pksc"), "rb").read()
    assert len(raw) == 8 + 28 * 3
    # halo i starts at byte 8 + 28 i (pksc.f90:64-66)
    assert struct.unpack("<f", raw[8 + 28 * 2 + 8:8 + 28 * 2 + 12])[0] == 3000.0


def test_pksc_round_trip_and_truncation(tmp_path):
    rng = np.random.default_rng(0)
    h = rng.normal(size=(1000, 7)).astype(np.float32)
    p = tmp_path / "c.pksc"
    formats.write_pksc(p, h)
    n, rmax, back = formats.read_pksc(p)
    assert n == 1000 and np.array_equal(back, h) and rmax == h[:, 6].max()
    with open(p, "r+b") as f:
        f.truncate(8 + 28 * 999 + 5)
    with pytest.raises(ValueError):
        formats.read_pksc(p)


def test_table_golden_fixture():
    hi, hf, t = formats.read_table(os.path.join(GOLD, "tiny_table.bin"))
    assert list(hi) == [3, 3, 4]
    assert np.array_equal(hf, np.array([1e2, 7e3, 1e13, 1e16, 1e-2, 10.0], np.float32))
    # column-major table(p, ir, imh, ichi): flat index (p-1) + 2((ir-1) + nrt((imh-1) + nmht(ichi-1)))
    for p in (1, 2):
        for ir in range(1, 5):
            for imh in range(1, 4):
                for ichi in range(1, 4):
                    idx = (p - 1) + 2 * ((ir - 1) + 4 * ((imh - 1) + 3 * (ichi - 1)))
                    assert t[idx] == 1000 * p + 100 * ir + 10 * imh + ichi


def test_table_round_trip(tmp_path):
    hi = np.array([5, 6, 7], np.int32)
    hf = np.array([1e2, 7e3, 1e13, 1e16, 1e-2, 10], np.float32)
    t = np.arange(2 * 5 * 6 * 7, dtype=np.float32)
    p = tmp_path / "t.bin"
    formats.write_table(p, hi, hf, t)
    assert os.path.getsize(p) == 36 + 4 * t.size
    a, b, c = formats.read_table(p)
    assert np.array_equal(a, hi) and np.array_equal(b, hf) and np.array_equal(c, t)


def test_fsmap_and_dpdy_layouts(tmp_path):
    m = np.arange(16, dtype=np.float32)
    p = tmp_path / "x_fs.map"
    formats.write_fsmap(p, 4, 0.1745329, m)
    raw = open(p, "rb").read()
    assert struct.unpack("<iiff", raw[:16])[:2] == (4, 4)
    assert len(raw) == 16 + 64
    n, fov, back = formats.read_fsmap(p)
    assert n == 4 and np.array_equal(back, m) and fov == pytest.approx(0.1745329)
    d = np.linspace(0, 1, 10000).astype(np.float32)
    q = tmp_path / "x_dp.bin"
    formats.write_dpdy(q, d)
    raw = open(q, "rb").read()
    assert struct.unpack("<if", raw[:8]) == (10000, np.float32(1e-4))
    assert len(raw) == 8 + 40000


@pytest.mark.parametrize("nside", [1, 4, 64])
def test_healpix_fits_round_trip(tmp_path, nside):
    npix = 12 * nside * nside
    m = np.random.default_rng(1).normal(size=npix).astype(np.float32)
    p = tmp_path / "m_hp.fits"
    formats.write_healpix_fits(p, m, nside)
    assert os.path.getsize(p) % 2880 == 0
    ns, order, back, hdus = formats.read_healpix_fits(p)
    assert ns == nside and order == "RING"
    assert np.array_equal(back[:npix], m)
    cards = hdus[1][0]
    rep = formats.healpix_repeat(npix)
    assert cards["TFORM1"] == f"{rep}E"
    assert int(cards["NAXIS2"]) == npix // rep
    assert cards["TTYPE1"] == "simulation"
    assert hdus[0][0]["SIMPLE"] == "T"
