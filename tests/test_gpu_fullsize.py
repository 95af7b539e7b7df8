"""Full-size workloads, checked through properties that do not need the
oracle to finish at that size:

  * C4 (10^8 halos, nside 4096, streamed in 2^24-halo chunks): the FP32 and
    FP64 tiled painters paint every pixel with exactly the same set of halos
    (per-pixel count and hash equal), and the FP32 map is within 1e-6 of the
    FP64 map on every non-zero pixel, with no pixel zero in only one of them;
  * C5 (10^8 halos, nside 8192, six frequency tables in one pass): the same
    1e-6 bound and zero pattern, map by map.

The FP64 tiled painter's own parity with the oracle (1e-12, bit-exact
membership) is tested at sizes the oracle finishes in seconds.
"""
import numpy as np
import pytest

from matt_young_mapmaking_amd import abi, catalog, devmem

pytestmark = pytest.mark.gpu

TABLES_C5 = [(148.0, 0.0), (200.0, 0.0), (201.0, 0.0), (218.0, 0.0), (353.0, 0.1), (545.0, 0.0)]


def paint(dev, n, nside, tabs, prec, membership, tile, chunk=0):
    p = abi.Painter(nside=nside, npix_fs=1024, precision=prec, n_maps=len(tabs), membership=membership,
                    chunk_halos=chunk)
    p.set_tile(*tile)
    p.set_run()
    for k, t in enumerate(tabs):
        p.load_table(k, *t)
    p.paint_device(dev.address, n)
    devmem.synchronize()
    return p


def compare(p32, p64, k):
    a = p32.maps(slot=k, dtype=np.float32)[0]
    b = p64.maps(slot=k, dtype=np.float64)[0]
    nz = b != 0
    rel = float(np.max(np.abs(a[nz].astype(np.float64) - b[nz]) / np.abs(b[nz])))
    zm = int(np.count_nonzero((a != 0) != nz))
    return rel, zm


@pytest.mark.parametrize("wl", ["C4", "C5"])
def test_fullsize_fp32_against_fp64(wl):
    cfg = catalog.CONFIGS[wl]
    n, nside = cfg["n"], cfg["nside"]
    print(f"\n[gpu fullsize] {wl}: {n} halos, nside {nside}", flush=True)
    h = catalog.config_catalog(wl)
    dev = devmem.DeviceArray(h)
    del h
    freqs = TABLES_C5 if wl == "C5" else [(148.0, 0.0)]
    tabs = [abi.make_maptable(model=1, nu=nu, alpha=a) for nu, a in freqs]
    tile = (1024, 256) if len(tabs) == 1 else (512, 256)
    member = wl == "C4"
    chunk = (1 << 24) if wl == "C4" else 0       # C4 is defined as 2^24-halo batches
    p32 = paint(dev, n, nside, tabs, abi.PREC_FP32_COMP, member, tile, chunk)
    print(f"[gpu fullsize] {wl} fp32 painted", flush=True)
    p64 = paint(dev, n, nside, tabs, abi.PREC_FP64, member, tile, chunk)
    print(f"[gpu fullsize] {wl} fp64 painted", flush=True)
    s32, s64 = p32.stats(), p64.stats()
    assert s32["n_upd_hp"] == s64["n_upd_hp"] and s32["n_painted"] == s64["n_painted"]
    assert s32["n_ring_gaps"] == 0 and s32["n_radius_err"] == 0
    if member:
        c32, h32, _, _ = p32.membership()
        c64, h64, _, _ = p64.membership()
        assert np.array_equal(c32, c64) and np.array_equal(h32, h64)
        assert int(c32.sum(dtype=np.uint64)) == s32["n_upd_hp"]
    for k in range(len(tabs)):
        rel, zm = compare(p32, p64, k)
        print(f"[gpu fullsize] {wl} map {k}: max rel err {rel:.3e}, zero mismatch {zm}", flush=True)
        assert rel <= 1e-6 and zm == 0
    p32.close()
    p64.close()
