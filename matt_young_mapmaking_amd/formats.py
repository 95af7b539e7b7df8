"""Byte layouts of every file pks2map reads or writes (native little-endian
stream I/O, the Intel `form='binary'` layout).

  .pksc     int32 N, float32 RTHmax, N x {x,y,z,vx,vy,vz,Rth} float32
            (writer merge_pkvd.f90:1311-1325, reader pksc.f90:46-68)
  table     int32 nchit,nmht,nrt; float32 chimint,chimaxt,mhmint,mhmaxt,rmint,rmaxt;
            float32 table(2,nrt,nmht,nchit) column-major (maptable.f90:94,142,158,162)
  _fs.map   int32 npix,npix; float32 fov,fov; float32 map(npix*npix) (pks2map.f90:372-375)
  _dp.bin   int32 ndpdy; float32 dpdy_max; float32 dpdy(ndpdy)   (pks2map.f90:366-369)
  _hp.fits  primary HDU + BINTABLE, one column 'simulation' of 1024E, NSIDE,
            ORDERING='RING' cards (fits_s_inc.f90:787-960, pks2map.f90:378-381)

The Fortran host (pks2map executable) writes the same files through the
library's writers; these numpy versions serve Python users and the tests.
"""
import datetime
import os

import numpy as np

PKSC_HDR = np.dtype([("n", "<i4"), ("rthmax", "<f4")])


def write_pksc(path, halos, rthmax=None):
    halos = np.ascontiguousarray(halos, dtype="<f4").reshape(-1, 7)
    if rthmax is None:
        rthmax = float(halos[:, 6].max()) if halos.shape[0] else 0.0
    with open(path, "wb") as f:
        np.array([(halos.shape[0], rthmax)], dtype=PKSC_HDR).tofile(f)
        halos.tofile(f)


def read_pksc(path, mmap=False):
    hdr = np.fromfile(path, dtype=PKSC_HDR, count=1)[0]
    n = int(hdr["n"])
    if mmap:
        h = np.memmap(path, dtype="<f4", mode="r", offset=8, shape=(n, 7))
    else:
        h = np.fromfile(path, dtype="<f4", offset=8, count=7 * n).reshape(n, 7)
    if h.shape[0] != n:
        raise ValueError(f"{path}: header says {n} halos, file holds {h.shape[0]}")
    return n, float(hdr["rthmax"]), h


def write_table(path, hdr_i, hdr_f, table):
    with open(path, "wb") as f:
        np.asarray(hdr_i, dtype="<i4").tofile(f)
        np.asarray(hdr_f, dtype="<f4").tofile(f)
        np.asarray(table, dtype="<f4").tofile(f)


def read_table(path):
    hdr_i = np.fromfile(path, dtype="<i4", count=3)
    hdr_f = np.fromfile(path, dtype="<f4", count=6, offset=12)
    n = 2 * int(hdr_i[0]) * int(hdr_i[1]) * int(hdr_i[2])
    t = np.fromfile(path, dtype="<f4", offset=36, count=n)
    if t.size != n:
        raise ValueError(f"{path}: header needs {n} floats, file holds {t.size}")
    return hdr_i, hdr_f, t


def write_fsmap(path, npix, fov_rad, m):
    with open(path, "wb") as f:
        np.array([npix, npix], "<i4").tofile(f)
        np.array([fov_rad, fov_rad], "<f4").tofile(f)
        np.asarray(m, "<f4").tofile(f)


def read_fsmap(path):
    n = np.fromfile(path, "<i4", count=2)
    fov = np.fromfile(path, "<f4", count=2, offset=8)
    m = np.fromfile(path, "<f4", offset=16)
    return int(n[0]), float(fov[0]), m


def write_dpdy(path, dpdy, dpdy_max=1e-4):
    with open(path, "wb") as f:
        np.array([len(dpdy)], "<i4").tofile(f)
        np.array([dpdy_max], "<f4").tofile(f)
        np.asarray(dpdy, "<f4").tofile(f)


def read_dpdy(path):
    n = int(np.fromfile(path, "<i4", count=1)[0])
    mx = float(np.fromfile(path, "<f4", count=1, offset=4)[0])
    d = np.fromfile(path, "<f4", offset=8, count=n)
    return n, mx, d


# ---------------------------------------------------------------- FITS
def _card(key, value=None, comment=None):
    if value is None:
        s = key.ljust(80)
    else:
        if isinstance(value, bool):
            v = ("T" if value else "F").rjust(20)
        elif isinstance(value, (int, np.integer)):
            v = str(int(value)).rjust(20)
        elif isinstance(value, str):
            v = ("'" + value.ljust(8) + "'").ljust(20)
        else:
            v = repr(float(value)).upper().rjust(20)
        s = f"{key:<8}= {v}"
        if comment:
            s += " / " + comment
    return s[:80].ljust(80)


def _block(cards):
    txt = "".join(cards) + "END".ljust(80)
    pad = (-len(txt)) % 2880
    return (txt + " " * pad).encode("ascii")


def healpix_repeat(npix):
    """Vector repeat count of the 'simulation' column (fits_s_inc.f90:888-893)."""
    repeat = 1024
    if npix < 1024:
        repeat = 1
    nside = int(round((npix / 12) ** 0.5))
    if nside > 1024 * 256:
        repeat = nside // 256
    return repeat


def write_healpix_fits(path, m, nside, date=None):
    """HEALPix RING map as a FITS BINTABLE with one float32 vector column."""
    m = np.asarray(m, dtype=np.float32).ravel()
    npix = m.size
    repeat = healpix_repeat(npix)
    nrows = npix // repeat
    date = date or datetime.date.today().isoformat()
    prim = [_card("SIMPLE", True, "file does conform to FITS standard"),
            _card("BITPIX", 32, "number of bits per data pixel"),
            _card("NAXIS", 0, "number of data axes"),
            _card("EXTEND", True, "FITS dataset may contain extensions"),
            _card("DATE", date, "file creation date (YYYY-MM-DD)")]
    ext = [_card("XTENSION", "BINTABLE", "binary table extension"),
           _card("BITPIX", 8, "8-bit bytes"),
           _card("NAXIS", 2, "2-dimensional binary table"),
           _card("NAXIS1", 4 * repeat, "width of table in bytes"),
           _card("NAXIS2", nrows, "number of rows in table"),
           _card("PCOUNT", 0, "size of special data area"),
           _card("GCOUNT", 1, "one data group (required keyword)"),
           _card("TFIELDS", 1, "number of fields in each row"),
           _card("TTYPE1", "simulation", "label for field   1"),
           _card("TFORM1", f"{repeat}E", "data format of field: 4-byte REAL"),
           _card("NSIDE", int(nside), "the nside of the map"),
           _card("ORDERING", "RING", "the nside of the map")]
    with open(path, "wb") as f:
        f.write(_block(prim))
        f.write(_block(ext))
        data = m.astype(">f4").tobytes()
        f.write(data)
        f.write(b"\0" * ((-len(data)) % 2880))


def read_healpix_fits(path):
    """Minimal reader for files written by write_healpix_fits / the library."""
    with open(path, "rb") as f:
        raw = f.read()
    pos = 0
    hdus = []
    while pos < len(raw):
        cards = {}
        while True:
            blk = raw[pos:pos + 2880].decode("ascii")
            pos += 2880
            done = False
            for i in range(36):
                c = blk[80 * i:80 * i + 80]
                k = c[:8].strip()
                if k == "END":
                    done = True
                    break
                if c[8:10] == "= ":
                    v = c[10:].split("/")[0].strip()
                    cards[k] = v.strip("'").strip()
            if done:
                break
        naxis = int(cards.get("NAXIS", 0))
        size = 0
        if naxis:
            size = int(cards.get("BITPIX", 8)) // 8 * np.prod([int(cards[f"NAXIS{i}"]) for i in range(1, naxis + 1)])
        data = raw[pos:pos + size]
        pos += size + ((-size) % 2880)
        hdus.append((cards, data))
        if pos >= len(raw):
            break
    cards, data = hdus[1]
    m = np.frombuffer(data, dtype=">f4").astype(np.float32)
    return int(cards["NSIDE"]), cards["ORDERING"], m, hdus


def file_size_ok(path, expected):
    return os.path.exists(path) and os.path.getsize(path) == expected
