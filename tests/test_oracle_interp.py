"""Pins of the oracle's interpolate_table (maptable.f90:176-216).

Synthetic tables make the interpolant's algebra visible:
  * a table linear in the indices, T = a*ir + b*imh + c*ichi + d*p, must give
    a*(ir+fr) + b*(imh+fc) + c*(ichi+fm) + d*p with fr, fc, fm the
    UN-normalised log offsets (quirk Q5) and the chi fraction on the mass axis
    and vice versa (quirk Q4) -- a textbook trilinear interpolation fails;
  * nodes are reproduced exactly; a constant table gives the constant;
  * r > rmaxt or chi > chimaxt gives 0; clamps of r and mh; the upper-edge
    index aliasing of Q6 reads the element of the flat column-major index.
"""
import math

import numpy as np
import pytest

import p2mo

NCHI, NM, NR = 5, 6, 7
HF = np.array([1e2, 7e3, 1e13, 1e16, 1e-2, 10], np.float32)
HI = np.array([NCHI, NM, NR], np.int32)


def make_table(fn):
    t = np.zeros((NCHI, NM, NR, 2), np.float32)   # [ichi][imh][ir][p], i.e. table(p,ir,imh,ichi)
    for ic in range(NCHI):
        for im in range(NM):
            for ir in range(NR):
                for p in range(2):
                    t[ic, im, ir, p] = fn(p + 1, ir + 1, im + 1, ic + 1)
    return t.ravel()


A, B, Cc, D = 3.0, 5.0, 7.0, 11.0
LIN = make_table(lambda p, ir, im, ic: A * ir + B * im + Cc * ic + D * p)


@pytest.fixture(scope="module")
def o():
    return p2mo.Oracle(HI, HF, LIN)


def grid(lo, hi, n):
    l0, l1 = math.log(float(np.float32(lo))), math.log(float(np.float32(hi)))
    return l0, (l1 - l0) / (n - 1)


def cell(x, lo, hi, n):
    l0, d = grid(lo, hi, n)
    u = (math.log(x) - l0) / d
    i = int(u) + 1
    return i, math.log(x) - (l0 + (i - 1) * d)


def test_linear_table_swapped_unnormalised_weights(o):
    rng = np.random.default_rng(1)
    worst_textbook = 0.0
    for _ in range(200):
        z = rng.uniform(0.1, 1.0)
        chi = o.rofz(z)
        mh = math.exp(rng.uniform(math.log(2e13), math.log(5e15)))
        r = math.exp(rng.uniform(math.log(0.02), math.log(8.0)))
        theta = math.atan(r / chi)
        v, _ = o.interpolate(theta, z, mh, profile=1)
        rr = chi * math.tan(theta)
        ir, fr = cell(rr, 1e-2, 10, NR)
        ic, fc = cell(chi, 1e2, 7e3, NCHI)
        im, fm = cell(mh, 1e13, 1e16, NM)
        expected = A * (ir + fr) + B * (im + fc) + Cc * (ic + fm) + D * 1
        assert v == pytest.approx(expected, rel=1e-12)
        textbook = A * (ir + fr / grid(1e-2, 10, NR)[1]) + B * (im + fm / grid(1e13, 1e16, NM)[1]) \
            + Cc * (ic + fc / grid(1e2, 7e3, NCHI)[1]) + D
        worst_textbook = max(worst_textbook, abs(v - textbook))
    # a textbook trilinear interpolation (normalised, unswapped) is far off
    assert worst_textbook > 0.5


def test_constant_table_and_nodes():
    oc = p2mo.Oracle(HI, HF, np.full(2 * NCHI * NM * NR, 4.25, np.float32))
    z = 0.4
    v, _ = oc.interpolate(math.atan(0.5 / oc.rofz(z)), z, 3e14)
    assert v == pytest.approx(4.25, rel=1e-14)


def test_cutoffs_and_clamps(o):
    z = 0.5
    chi = o.rofz(z)
    # r > rmaxt -> 0
    assert o.interpolate(math.atan(10.5 / chi), z, 1e14)[0] == 0.0
    # chi > chimaxt -> 0 (rofz(5.9) exceeds 7000 Mpc)
    assert o.rofz(5.9) > 7000.0
    assert o.interpolate(1e-5, 5.9, 1e14)[0] == 0.0
    # r < rmint is clamped to rmint + 1e-5
    v_small, _ = o.interpolate(1e-9, z, 1e14)
    v_clamp, _ = o.interpolate(math.atan((float(np.float32(1e-2)) + 1e-5) / chi), z, 1e14)
    assert v_small == pytest.approx(v_clamp, rel=1e-9)
    # mh below the table is clamped and written back (quirk Q3)
    v_lo, m_out = o.interpolate(math.atan(1.0 / chi), z, 1e12)
    assert m_out == pytest.approx(float(np.float32(1e13)), rel=1e-15)
    v_at, _ = o.interpolate(math.atan(1.0 / chi), z, float(np.float32(1e13)))
    assert v_lo == v_at


def test_upper_edge_aliasing(o):
    """mh clamped to mhmaxt - 1e-5 (== mhmaxt) gives imh = nmht, so the imh+1
    corners read the next chi slab of the column-major array (quirk Q6)."""
    rng = np.random.default_rng(5)
    vals = rng.uniform(1, 2, 2 * NCHI * NM * NR).astype(np.float32)
    oa = p2mo.Oracle(HI, HF, vals)
    z = 0.3
    chi = oa.rofz(z)
    r = 0.5
    v, m_out = oa.interpolate(math.atan(r / chi), z, 5e16)
    assert m_out == float(np.float32(1e16))
    flat = np.concatenate([vals, np.zeros(10000, np.float32)])

    def T(p, ir, im, ic):
        return float(flat[(p - 1) + 2 * ((ir - 1) + NR * ((im - 1) + NM * (ic - 1)))])

    rr = chi * math.tan(math.atan(r / chi))
    ir, fr = cell(rr, 1e-2, 10, NR)
    ic, fc = cell(chi, 1e2, 7e3, NCHI)
    l0, dm = grid(1e13, 1e16, NM)
    lm = math.log(m_out)
    im = int((lm - l0) / dm) + 1
    fm = lm - (l0 + (im - 1) * dm)
    assert im == NM
    exp = 0.0
    for a, wa in ((0, 1 - fr), (1, fr)):
        for b, wb in ((0, 1 - fc), (1, fc)):
            for c, wc in ((0, 1 - fm), (1, fm)):
                exp += T(1, ir + a, im + b, ic + c) * wa * wb * wc
    assert v == pytest.approx(exp, rel=1e-12)
