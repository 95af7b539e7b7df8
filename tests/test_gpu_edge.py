"""Edge cases on the GPU: empty and all-cut catalogues, discs centred on the
poles (the a == 0 branch of discphirange_at_z), discs covering the pole or
most of the sky, halos whose disc radius is invalid (r >= chi, NaN), the kSZ
profile, and the drop-in Fortran executable end to end, with its scramble and
centring branches."""
import os
import subprocess

import numpy as np
import pytest

import p2mo
from matt_young_mapmaking_amd import abi, catalog, formats

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def table():
    return abi.make_maptable(model=1, nu=148.0, alpha=0.0)


def paint_both(h, table, nside=256, npix=128, prec=abi.PREC_FP64, profile=1):
    o = p2mo.Oracle(*table)
    ref = o.paint(h, nside=nside, npix=npix, nthreads=4, profile=profile)
    with abi.Painter(nside=nside, npix_fs=npix, precision=prec, membership=True, profile=profile) as p:
        p.set_run()
        p.load_table(0, *table)
        p.paint(h)
        hp, fs = p.maps()
        mem = p.membership()
        st = p.stats()
    return ref, hp, fs, mem, st


@pytest.mark.parametrize("prec", [abi.PREC_PLAIN_FP64, abi.PREC_FP64, abi.PREC_FP32_COMP])
def test_empty_and_all_cut(table, prec):
    print(f"\n[gpu edge] empty precision={prec}", flush=True)
    for h in (np.zeros((0, 7), np.float32), np.array([[0, 0, 50.0, 0, 0, 0, 5.0]], np.float32)):
        with abi.Painter(nside=64, npix_fs=32, precision=prec, membership=True) as p:
            p.set_run()
            p.load_table(0, *table)
            p.paint(h)
            hp, fs = p.maps()
            hc, hh, fc, fh = p.membership()
        assert not hp.any() and not fs.any() and not hc.any() and not fc.any()


@pytest.mark.parametrize("prec", [abi.PREC_PLAIN_FP64, abi.PREC_FP64, abi.PREC_FP32_COMP])
def test_poles_and_large_discs(table, prec):
    print(f"\n[gpu edge] poles precision={prec}", flush=True)
    h = np.array([
        [0.0, 0.0, 300.0, 0, 0, 0, 12.0],      # on the painter x axis (flat-sky centre)
        [300.0, 0.0, 0.0, 0, 0, 0, 12.0],      # hpvec z = pos1/chi = +1: north pole
        [-300.0, 0.0, 0.0, 0, 0, 0, 12.0],     # south pole
        [280.0, 30.0, 0.0, 0, 0, 0, 30.0],     # disc over the pole
        [220.0, 0.0, 0.0, 0, 0, 0, 60.0],      # very large disc
    ], np.float32)
    ref, hp, fs, (hc, hh, fc, fh), st = paint_both(h, table, nside=512, npix=128, prec=prec)
    assert np.array_equal(hc, ref["hp_count"]) and np.array_equal(hh, ref["hp_hash"])
    assert np.array_equal(fc, ref["fs_count"])
    nz = ref["hp"] != 0
    tol = 1e-6 if prec == abi.PREC_FP32_COMP else 1e-12
    assert np.max(np.abs(hp[nz] / ref["hp"][nz] - 1)) <= tol


@pytest.mark.parametrize("prec", [abi.PREC_FP64, abi.PREC_FP32_COMP])
def test_runs_over_many_tiles(table, prec):
    """Discs ~0.35 rad across on 64-pixel tiles at nside 2048: every ring run
    spans about 14 tiles, so most pieces come after their item's first and go
    through the binning's overflow list, whose first size (n_items / 8 +
    65536 entries) is too small here: the binning grows it and counts again."""
    print(f"\n[gpu edge] many-tile runs precision={prec}", flush=True)
    h = np.array([
        [0.0, 250.0, 0.0, 0, 0, 0, 200.0],
        [0.0, -250.0, 0.0, 0, 0, 0, 200.0],
        [0.0, 0.0, 250.0, 0, 0, 0, 200.0],
        [60.0, 0.0, -240.0, 0, 0, 0, 200.0],
        [0.0, 180.0, 180.0, 0, 0, 0, 120.0],
        [40.0, -180.0, 170.0, 0, 0, 0, 200.0],
        [-50.0, 170.0, -180.0, 0, 0, 0, 200.0],
        [-30.0, -170.0, -175.0, 0, 0, 0, 200.0],
    ], np.float32)
    o = p2mo.Oracle(*table)
    ref = o.paint(h, nside=2048, npix=64, nthreads=8)
    with abi.Painter(nside=2048, npix_fs=64, precision=prec, membership=True) as p:
        p.set_tile(64, 64)
        p.set_run()
        p.load_table(0, *table)
        p.paint(h)
        hp = p.maps()[0]
        hc, hh, _, _ = p.membership()
        st = p.stats()
    assert st["n_segments"] > 140_000
    assert np.array_equal(hc, ref["hp_count"]) and np.array_equal(hh, ref["hp_hash"])
    nz = ref["hp"] != 0
    assert nz.sum() > 50_000 and hc.sum() > 5_000_000
    tol = 1e-6 if prec == abi.PREC_FP32_COMP else 1e-12
    assert np.max(np.abs(hp[nz] / ref["hp"][nz] - 1)) <= tol


def test_invalid_radius_is_counted_and_skipped(table):
    # rth so large that rh > chi: asin(rh/chi) is undefined (quirk Q16)
    h = np.array([[0, 0, 220.0, 0, 0, 0, 5000.0],
                  [0, 0, 1500.0, 0, 0, 0, 8.0],
                  [np.nan, 0, 1500.0, 0, 0, 0, 8.0]], np.float32)
    with abi.Painter(nside=256, npix_fs=64, precision=abi.PREC_FP32_COMP, membership=True) as p:
        p.set_run()
        p.load_table(0, *table)
        p.paint(h)
        st = p.stats()
        hc = p.membership()[0]
    assert st["n_radius_err"] == 2
    assert st["n_painted"] == 1
    o = p2mo.Oracle(*table)
    ref = o.paint(h[1:2], nside=256, npix=64)
    assert np.array_equal(hc, ref["hp_count"])


def test_ksz_profile(table):
    h = catalog.config_catalog("C2", n=20000)
    ref, hp, fs, (hc, hh, fc, fh), st = paint_both(h, table, nside=512, npix=128, prec=abi.PREC_FP64, profile=2)
    assert np.array_equal(hh, ref["hp_hash"])
    nz = ref["hp"] != 0
    assert np.max(np.abs(hp[nz] / ref["hp"][nz] - 1)) <= 1e-12
    assert (hp > 0).any() and (hp < 0).any()      # -v_r/c takes both signs


def test_fortran_pks2map_end_to_end(table, tmp_path):
    exe = os.path.join(ROOT, "matt_young_mapmaking_amd", "bin", "pks2map")
    if not os.path.exists(exe):
        pytest.skip("Fortran host not built")
    h = catalog.config_catalog("C1")
    formats.write_pksc(tmp_path / "cat.pksc", h)
    formats.write_table(tmp_path / "tab_148", *table)
    r = subprocess.run([exe, "cat.pksc", "out", "tab_148", "0.05", "1024", "0", "0", "512", "10", "1.2", "0", "1", "1"],
                       cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert "Number of halos in map:    17299" in r.stdout
    assert "Healpix moments:" in r.stdout and "Flatsky RMS:" in r.stdout
    # the executable parses CLI reals as float32, as the reference does
    o = p2mo.Oracle(*table)
    ref = o.paint(h, nside=1024, npix=512, zmin=float(np.float32(0.05)), zmax=float(np.float32(1.2)),
                  fov_deg=10.0, nthreads=8)
    ns, order, hp, _ = formats.read_healpix_fits(str(tmp_path / "out_hp.fits"))
    nz = ref["hp"] != 0
    assert np.count_nonzero((hp != 0) != nz) == 0
    assert np.max(np.abs(hp[nz] / ref["hp"][nz] - 1)) < 1e-6
    n, fov, fs = formats.read_fsmap(str(tmp_path / "out_fs.map"))
    nzf = ref["fs"] != 0
    assert np.max(np.abs(fs[nzf] / ref["fs"][nzf] - 1)) < 1e-6
    assert os.path.getsize(tmp_path / "out_dp.bin") == 8 + 40000


@pytest.mark.parametrize("scramble,centre", [(1, 0), (0, 1), (1, 1)])
def test_fortran_pks2map_scramble_and_centre(table, tmp_path, scramble, centre):
    """The host branches of the drop-in (pks2map.f90:162-262): the records it
    paints equal the oracle's restatement of the branches to float rounding,
    and its map equals the oracle's map of those records."""
    print(f"\n[gpu edge] pks2map scramble={scramble} centre={centre}", flush=True)
    exe = os.path.join(ROOT, "matt_young_mapmaking_amd", "bin", "pks2map")
    if not os.path.exists(exe):
        pytest.skip("Fortran host not built")
    h = catalog.config_catalog("C2", n=40000)
    formats.write_pksc(tmp_path / "cat.pksc", h)
    formats.write_table(tmp_path / "tab_148", *table)
    env = dict(os.environ, P2M_DUMP_CATALOGUE=str(tmp_path / "painted.pksc"))
    r = subprocess.run([exe, "cat.pksc", "out", "tab_148", "0.05", "512", str(scramble), str(centre), "128", "10",
                        "1.2", "0", "1", "1"], cwd=tmp_path, capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stderr
    _, _, painted = formats.read_pksc(str(tmp_path / "painted.pksc"))
    o = p2mo.Oracle(*table)
    lo, hi = o.rofz(float(np.float32(0.05))), o.rofz(float(np.float32(1.2)))
    expect = p2mo.host_branches(h, np.float32(lo), np.float32(hi), 0.0, scramble=scramble, centre=centre)
    assert painted.shape == expect.shape
    fin = np.all(np.isfinite(expect[:, :3]), axis=1)
    assert np.array_equal(fin, np.all(np.isfinite(painted[:, :3]), axis=1))
    assert (not scramble) or not fin[0]                  # Q10: the first scrambled halo is NaN
    assert fin.sum() > 30000
    chi = np.linalg.norm(expect[fin, :3].astype(np.float64), axis=1)
    dev = np.max(np.abs(painted[fin, :3].astype(np.float64) - expect[fin, :3]), axis=1) / chi
    assert np.max(dev) < 1e-5
    assert np.array_equal(painted[:, 6], expect[:, 6])
    # the map of the painted records (radial cut already applied: the band is widened so
    # that float rounding of the rotated distances cannot drop a record)
    ref = o.paint(painted[fin], nside=512, npix=128, zmin=0.04, zmax=1.3, fov_deg=10.0, nthreads=8)
    ns, order, hp, _ = formats.read_healpix_fits(str(tmp_path / "out_hp.fits"))
    nz = ref["hp"] != 0
    assert np.count_nonzero((hp != 0) != nz) == 0
    assert np.max(np.abs(hp[nz] / ref["hp"][nz] - 1)) < 1e-6


def test_clamped_radius_cell_of_a_fine_radial_grid():
    """A table whose radial step is smaller than ln(1 + 1e-5 / rmint): the
    clamp r = rmint + 1e-5 (maptable.f90:190) lands in cell 2, not 1.  Nearly
    every painted pixel has r < rmint here, so the FP32 painter must take the
    clamped cell and fraction the reference forms (:196, :200); it matches the
    oracle within 1e-6 with bit-exact membership."""
    print("\n[gpu edge] clamped radius on a fine radial grid", flush=True)
    nchit, nmht, nrt = 20, 10, 200
    hi = np.array([nchit, nmht, nrt], np.int32)
    hf = np.array([1e2, 7e3, 1e13, 1e16, 1.0, 1.00199], np.float32)
    drt = (np.log(np.float64(hf[5])) - np.log(np.float64(hf[4]))) / (nrt - 1)
    assert int(np.log(np.float64(hf[4]) + 1e-5) - np.log(np.float64(hf[4])) >= drt) == 1   # cell 2
    ir = np.arange(1, nrt + 1, dtype=np.float64)
    t = np.zeros((nchit, nmht, nrt, 2), np.float32)
    for ichi in range(nchit):
        for imh in range(nmht):
            t[ichi, imh, :, 0] = -(1.0 + 0.5 * ir + 0.01 * imh + 0.003 * ichi) * 1e-7
            t[ichi, imh, :, 1] = 1e-4
    tab = (hi, hf, t.ravel())
    h = catalog.config_catalog("C2", n=20000)
    ref, hp, fs, (hc, hh, fc, fh), st = paint_both(h, tab, nside=1024, npix=256, prec=abi.PREC_FP32_COMP)
    assert np.array_equal(hc, ref["hp_count"]) and np.array_equal(hh, ref["hp_hash"])
    nz = ref["hp"] != 0
    assert nz.sum() > 1000
    assert np.max(np.abs(hp[nz] / ref["hp"][nz] - 1)) <= 1e-6


@pytest.mark.parametrize("prec", [abi.PREC_FP64, abi.PREC_FP32_COMP])
def test_dense_tiles_take_several_batches(table, prec):
    """4000 discs clustered within about 0.03 rad of one equatorial direction
    at nside 512 (up to 1200 halos on one pixel): the tiles under the cluster
    list up to 2e4 pixels, far more than one batch holds (3072), and hundreds
    of segments, so the painters stage them in several batches per tile
    (later batches behind their own barrier, the segment-written byte map
    per batch, the batch counts formed anew).  Membership and values against
    the oracle."""
    print(f"\n[gpu edge] dense tiles precision={prec}", flush=True)
    rng = np.random.default_rng(2468)
    n = 4000
    h = np.zeros((n, 7), np.float32)
    c0 = 250.0 / np.sqrt(2.0)                            # chi = 250 Mpc, beyond the z = 0.05 cut
    h[:, 0] = rng.uniform(-5.0, 5.0, n)                  # hpvec z = pos[0] / chi: near the equator
    h[:, 1] = c0 + rng.uniform(-5.0, 5.0, n)
    h[:, 2] = c0 + rng.uniform(-5.0, 5.0, n)
    h[:, 6] = rng.uniform(5.0, 9.0, n)                   # Lagrangian radius in Mpc
    ref, hp, fs, (hc, hh, fc, fh), st = paint_both(h, table, nside=512, npix=64, prec=prec)
    # in the equatorial belt every ring starts on a multiple of 1024 and holds
    # 2048 pixels, so the 1024-pixel tiles are the aligned 1024-pixel windows
    per_tile = ref["hp_count"].astype(np.int64).reshape(-1, 1024).sum(axis=1)
    assert per_tile.max() > 3 * 3072, per_tile.max()
    assert np.array_equal(hc, ref["hp_count"]) and np.array_equal(hh, ref["hp_hash"])
    nz = ref["hp"] != 0
    assert np.array_equal(hp != 0, nz)
    tol = 1e-6 if prec == abi.PREC_FP32_COMP else 1e-12
    assert np.max(np.abs(hp[nz] / ref["hp"][nz] - 1)) <= tol
