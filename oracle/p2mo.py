"""ctypes front end of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as
the checker of the HIP library; never by the library or its executables.
Builds with `make -C oracle` (also done by __graft_entry__.build()).
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS = {}


class Cosmo(C.Structure):
    _fields_ = [("omegam", C.c_double), ("omegal", C.c_double), ("h", C.c_double),
                ("rho_0", C.c_double), ("rhocrit_0", C.c_double),
                ("nrofz", C.c_int), ("nzofr", C.c_int),
                ("rofz_min", C.c_double), ("rofz_max", C.c_double), ("drofz", C.c_double),
                ("zofr_min", C.c_double), ("zofr_max", C.c_double), ("dzofr", C.c_double),
                ("rofz", C.POINTER(C.c_double)), ("zofr", C.POINTER(C.c_double))]


class Table(C.Structure):
    _fields_ = [("nchit", C.c_int32), ("nmht", C.c_int32), ("nrt", C.c_int32),
                ("chimint", C.c_double), ("chimaxt", C.c_double), ("mhmint", C.c_double),
                ("mhmaxt", C.c_double), ("rmint", C.c_double), ("rmaxt", C.c_double),
                ("drt", C.c_double), ("dchit", C.c_double), ("dmht", C.c_double),
                ("t", C.POINTER(C.c_float)), ("nt", C.c_int64)]


class Run(C.Structure):
    _fields_ = [("nside", C.c_int32), ("npix", C.c_int32), ("fov_deg", C.c_double),
                ("zmin", C.c_double), ("zmax", C.c_double), ("chihview", C.c_double),
                ("model", C.c_int32), ("profile", C.c_int32), ("mmin", C.c_double),
                ("nthreads", C.c_int32)]


class Stats(C.Structure):
    _fields_ = [("n_in", C.c_int64), ("n_cut", C.c_int64), ("n_painted", C.c_int64),
                ("n_listed_hp", C.c_int64), ("n_upd_hp", C.c_int64), ("n_upd_fs", C.c_int64),
                ("n_ring_gaps", C.c_int64), ("cut_low", C.c_double), ("cut_high", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def _ptr(a, ct):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(ct))


def lib(flavour="portable"):
    """Load the oracle shared library; flavour 'portable' or 'libm'."""
    name = "libp2mo.so" if flavour == "portable" else "libp2mo_libm.so"
    if name in _LIBS:
        return _LIBS[name]
    path = os.path.join(_HERE, "build", name)
    if not os.path.exists(path):
        raise RuntimeError(f"oracle library {path} is not built; run `make -C oracle`")
    L = C.CDLL(path)
    d, i32, i64 = C.c_double, C.c_int32, C.c_int64
    L.p2mo_cosmo_init.argtypes = [C.POINTER(Cosmo), d, d, d]
    L.p2mo_cosmo_free.argtypes = [C.POINTER(Cosmo)]
    for f in ("p2mo_zofr", "p2mo_rofz", "p2mo_rhocrit", "p2mo_deltacrit"):
        getattr(L, f).argtypes = [C.POINTER(Cosmo), d]
        getattr(L, f).restype = d
    L.p2mo_table_init.argtypes = [C.POINTER(Table), C.POINTER(i32), C.POINTER(C.c_float), C.POINTER(C.c_float), i64]
    L.p2mo_interpolate_table.argtypes = [C.POINTER(Table), C.POINTER(Cosmo), d, d, C.POINTER(d), C.c_int]
    L.p2mo_interpolate_table.restype = d
    L.p2mo_nside2npix.argtypes = [i32]
    L.p2mo_nside2npix.restype = i64
    L.p2mo_ring_num.argtypes = [i32, d]
    L.p2mo_ring_num.restype = i32
    L.p2mo_ring2z.argtypes = [i32, i32]
    L.p2mo_ring2z.restype = d
    L.p2mo_pix2vec_ring.argtypes = [i32, i64, C.POINTER(d)]
    L.p2mo_angdist.argtypes = [C.POINTER(d), C.POINTER(d)]
    L.p2mo_angdist.restype = d
    L.p2mo_query_disc.argtypes = [i32, C.POINTER(d), d, C.POINTER(i32), C.POINTER(i64)]
    L.p2mo_query_disc.restype = i64
    L.p2mo_paint.argtypes = [C.POINTER(Run), C.POINTER(Cosmo), C.POINTER(Table), C.POINTER(C.c_float), i64,
                             C.POINTER(d), C.POINTER(d), C.POINTER(C.c_uint32), C.POINTER(C.c_uint64),
                             C.POINTER(C.c_uint32), C.POINTER(C.c_uint64), C.POINTER(Stats)]
    L.p2mo_halo_hash.argtypes = [i64]
    L.p2mo_halo_hash.restype = C.c_uint64
    L.p2mo_make_maptable.argtypes = [i32, d, d, i32, i32, i32, i64, i64, C.POINTER(i32), C.POINTER(C.c_float),
                                     C.POINTER(C.c_float), i32]
    L.p2mo_math.argtypes = [C.c_int, C.POINTER(d), C.POINTER(d), C.POINTER(d), i64]
    L.p2mo_ranf.argtypes = [C.POINTER(i32)]
    L.p2mo_ranf.restype = d
    L.p2mo_host_branches.argtypes = [C.POINTER(C.c_float), C.POINTER(C.c_int64), C.c_float, C.c_float, C.c_float,
                                     C.c_int, C.c_int]
    L.p2mo_host_branches.restype = None
    L.p2mo_halo_setup.argtypes = [C.POINTER(Run), C.POINTER(Cosmo), C.POINTER(C.c_float), i64, C.POINTER(d)]
    L.p2mo_paint_ref32.argtypes = [C.POINTER(Run), C.POINTER(i32), C.POINTER(C.c_float), C.POINTER(C.c_float), i64,
                                   C.POINTER(C.c_float), i64, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                   C.POINTER(C.c_uint32), C.POINTER(C.c_uint64), C.POINTER(C.c_uint32),
                                   C.POINTER(C.c_uint64), C.POINTER(d), C.POINTER(Stats)]
    L.p2mo_ref32_tables.argtypes = [d, d, d, C.POINTER(C.c_float), C.POINTER(C.c_float)]
    L.p2mo_ref32_interpolate.argtypes = [C.POINTER(i32), C.POINTER(C.c_float), C.POINTER(C.c_float), i64, d, d, d,
                                         C.c_float, C.c_float, C.POINTER(C.c_float), C.c_int]
    L.p2mo_ref32_interpolate.restype = C.c_float
    _LIBS[name] = L
    return L


# pks2map.f90:118-121 hard-coded cosmology
PKS2MAP_COSMO = dict(omegam=0.25, omegal=0.75, h=0.7)


class Oracle:
    """Cosmology + table state of one pks2map run, in the oracle."""

    def __init__(self, hdr_i, hdr_f, table, flavour="portable", omegam=0.25, omegal=0.75, h=0.7):
        self.L = lib(flavour)
        self.cosmo = Cosmo()
        self.L.p2mo_cosmo_init(C.byref(self.cosmo), omegam, omegal, h)
        self.hdr_i = np.ascontiguousarray(hdr_i, dtype=np.int32)
        self.hdr_f = np.ascontiguousarray(hdr_f, dtype=np.float32)
        self.table_padded = pad_table(table)
        self.table = Table()
        self.L.p2mo_table_init(C.byref(self.table), _ptr(self.hdr_i, C.c_int32), _ptr(self.hdr_f, C.c_float),
                               _ptr(self.table_padded, C.c_float), self.table_padded.size)

    def __del__(self):
        try:
            self.L.p2mo_cosmo_free(C.byref(self.cosmo))
        except Exception:
            pass

    def zofr(self, r):
        return self.L.p2mo_zofr(C.byref(self.cosmo), float(r))

    def rofz(self, z):
        return self.L.p2mo_rofz(C.byref(self.cosmo), float(z))

    def rhocrit(self, z):
        return self.L.p2mo_rhocrit(C.byref(self.cosmo), float(z))

    def deltacrit(self, z):
        return self.L.p2mo_deltacrit(C.byref(self.cosmo), float(z))

    def interpolate(self, theta, z, mh, profile=1):
        m = C.c_double(mh)
        v = self.L.p2mo_interpolate_table(C.byref(self.table), C.byref(self.cosmo), float(theta), float(z), C.byref(m), profile)
        return v, m.value

    def paint(self, halos, nside=0, npix=0, fov_deg=10.0, zmin=0.05, zmax=1.2, chihview=0.0, model=1,
              profile=1, mmin=1e13, nthreads=1, maps=True, membership=True):
        halos = np.ascontiguousarray(halos, dtype=np.float32).reshape(-1, 7)
        run = Run(nside, npix, fov_deg, zmin, zmax, chihview, model, profile, mmin, nthreads)
        hpn = 12 * nside * nside if nside > 0 else 0
        out = {}
        hp = np.zeros(hpn, np.float64) if (maps and hpn) else None
        fs = np.zeros(npix * npix, np.float64) if (maps and npix) else None
        hc = np.zeros(hpn, np.uint32) if (membership and hpn) else None
        hh = np.zeros(hpn, np.uint64) if (membership and hpn) else None
        fc = np.zeros(npix * npix, np.uint32) if (membership and npix) else None
        fh = np.zeros(npix * npix, np.uint64) if (membership and npix) else None
        st = Stats()
        rc = self.L.p2mo_paint(C.byref(run), C.byref(self.cosmo), C.byref(self.table), _ptr(halos, C.c_float),
                               halos.shape[0], _ptr(hp, C.c_double), _ptr(fs, C.c_double),
                               _ptr(hc, C.c_uint32), _ptr(hh, C.c_uint64), _ptr(fc, C.c_uint32), _ptr(fh, C.c_uint64),
                               C.byref(st))
        if rc != 0:
            raise RuntimeError(f"p2mo_paint failed: {rc}")
        out.update(hp=hp, fs=fs, hp_count=hc, hp_hash=hh, fs_count=fc, fs_hash=fh, stats=st.as_dict())
        return out

    def halo_setup(self, halos, zmin=0.05, zmax=1.2, chihview=0.0, model=1, profile=1, mmin=1e13):
        """Per-halo (z, mh, rh, chi, dtheta, painted) of the FP64 halo loop."""
        halos = np.ascontiguousarray(halos, dtype=np.float32).reshape(-1, 7)
        run = Run(0, 0, 10.0, zmin, zmax, chihview, model, profile, mmin, 1)
        out = np.zeros((halos.shape[0], 6), np.float64)
        self.L.p2mo_halo_setup(C.byref(run), C.byref(self.cosmo), _ptr(halos, C.c_float), halos.shape[0],
                               _ptr(out, C.c_double))
        return out

    def paint_ref32(self, halos, nside=0, npix=0, fov_deg=10.0, zmin=0.05, zmax=1.2, chihview=0.0, model=1,
                    profile=1, mmin=1e13, nthreads=1, maps=True, membership=False, setup=False):
        """The halo loop in the reference's float32 arithmetic (ref32 mode):
        float32 maps summed from thread-private maps, optional membership and
        per-halo setup values (z, mh, rh, chi, dtheta, painted)."""
        halos = np.ascontiguousarray(halos, dtype=np.float32).reshape(-1, 7)
        run = Run(nside, npix, fov_deg, zmin, zmax, chihview, model, profile, mmin, nthreads)
        hpn = 12 * nside * nside if nside > 0 else 0
        hp = np.zeros(hpn, np.float32) if (maps and hpn) else None
        fs = np.zeros(npix * npix, np.float32) if (maps and npix) else None
        hc = np.zeros(hpn, np.uint32) if (membership and hpn) else None
        hh = np.zeros(hpn, np.uint64) if (membership and hpn) else None
        fc = np.zeros(npix * npix, np.uint32) if (membership and npix) else None
        fh = np.zeros(npix * npix, np.uint64) if (membership and npix) else None
        su = np.zeros((halos.shape[0], 6), np.float64) if setup else None
        st = Stats()
        rc = self.L.p2mo_paint_ref32(C.byref(run), _ptr(self.hdr_i, C.c_int32), _ptr(self.hdr_f, C.c_float),
                                     _ptr(self.table_padded, C.c_float), self.table_padded.size,
                                     _ptr(halos, C.c_float), halos.shape[0], _ptr(hp, C.c_float), _ptr(fs, C.c_float),
                                     _ptr(hc, C.c_uint32), _ptr(hh, C.c_uint64), _ptr(fc, C.c_uint32),
                                     _ptr(fh, C.c_uint64), _ptr(su, C.c_double), C.byref(st))
        if rc != 0:
            raise RuntimeError(f"p2mo_paint_ref32 failed: {rc}")
        return dict(hp=hp, fs=fs, hp_count=hc, hp_hash=hh, fs_count=fc, fs_hash=fh, setup=su, stats=st.as_dict())

    def interpolate_ref32(self, theta, z, mh, profile=1):
        m = C.c_float(mh)
        v = self.L.p2mo_ref32_interpolate(_ptr(self.hdr_i, C.c_int32), _ptr(self.hdr_f, C.c_float),
                                          _ptr(self.table_padded, C.c_float), self.table_padded.size,
                                          0.25, 0.75, 0.7, C.c_float(theta), C.c_float(z), C.byref(m), profile)
        return float(v), float(m.value)


def pad_table(table):
    """Table as a flat float32 array followed by zeros that cover every
    out-of-range corner interpolate_table can address (quirk Q6)."""
    t = np.ascontiguousarray(table, dtype=np.float32).ravel()
    return np.concatenate([t, np.zeros(4096 + t.size // 50, np.float32)])


def query_disc(nside, vec, radius, flavour="portable"):
    L = lib(flavour)
    v = np.ascontiguousarray(vec, dtype=np.float64)
    lst = np.zeros(12 * nside * nside, np.int32)
    gaps = C.c_int64(0)
    n = L.p2mo_query_disc(nside, _ptr(v, C.c_double), float(radius), _ptr(lst, C.c_int32), C.byref(gaps))
    if n < 0:
        raise ValueError("radius out of [0, pi]")
    return lst[:n].copy(), gaps.value


def pix2vec_ring(nside, ipix, flavour="portable"):
    L = lib(flavour)
    v = np.zeros(3, np.float64)
    L.p2mo_pix2vec_ring(nside, int(ipix), _ptr(v, C.c_double))
    return v


def angdist(v1, v2, flavour="portable"):
    L = lib(flavour)
    a = np.ascontiguousarray(v1, np.float64)
    b = np.ascontiguousarray(v2, np.float64)
    return L.p2mo_angdist(_ptr(a, C.c_double), _ptr(b, C.c_double))


def ring_num(nside, z, flavour="portable"):
    return lib(flavour).p2mo_ring_num(nside, float(z))


def ring2z(nside, ir, flavour="portable"):
    return lib(flavour).p2mo_ring2z(nside, int(ir))


def halo_hash(i, flavour="portable"):
    return lib(flavour).p2mo_halo_hash(int(i))


def make_maptable(model=1, nu=148.0, alpha=0.0, nchit=100, nmht=100, nrt=200, nodes=None, nthreads=8,
                  flavour="portable"):
    """Returns (hdr_i, hdr_f, table[2*nrt*nmht*nchit]).  `nodes` = (lo, hi)
    restricts the computed (chi, mass) nodes; the rest stay zero."""
    L = lib(flavour)
    hdr_i = np.zeros(3, np.int32)
    hdr_f = np.zeros(6, np.float32)
    tab = np.zeros(2 * nrt * nmht * nchit, np.float32)
    lo, hi = nodes if nodes is not None else (0, nchit * nmht)
    L.p2mo_make_maptable(model, nu, alpha, nchit, nmht, nrt, lo, hi, _ptr(hdr_i, C.c_int32),
                         _ptr(hdr_f, C.c_float), _ptr(tab, C.c_float), nthreads)
    return hdr_i, hdr_f, tab


def host_branches(halos, cut_low, cut_high, chihview=0.0, scramble=0, centre=0):
    """Radial cut, scramble and centring of pks2map.f90:144-262 on a copy of
    the (n, 7) float32 records; returns the kept records."""
    h = np.ascontiguousarray(halos, np.float32).copy()
    n = C.c_int64(h.shape[0])
    lib().p2mo_host_branches(_ptr(h, C.c_float), C.byref(n), C.c_float(cut_low), C.c_float(cut_high),
                             C.c_float(chihview), int(scramble), int(centre))
    return h[:n.value].copy()


def ref32_tables(omegam=0.25, omegal=0.75, h=0.7):
    """The float32 rofz and zofr tables of cosmology.f90:51-105."""
    rofz = np.zeros(10000, np.float32)
    zofr = np.zeros(10000, np.float32)
    lib().p2mo_ref32_tables(omegam, omegal, h, _ptr(rofz, C.c_float), _ptr(zofr, C.c_float))
    return rofz, zofr


def ranf(seed):
    s = (C.c_int32 * 4)(*seed)
    r = lib().p2mo_ranf(s)
    return r, list(s)


MATH_OPS = ["sqrt", "sin", "cos", "tan", "asin", "acos", "atan2", "log", "exp", "pow"]


def math(op, x, y=None, flavour="portable"):
    L = lib(flavour)
    x = np.ascontiguousarray(x, np.float64)
    yy = None if y is None else np.ascontiguousarray(y, np.float64)
    out = np.zeros_like(x)
    rc = L.p2mo_math(MATH_OPS.index(op), _ptr(x, C.c_double), _ptr(yy, C.c_double), _ptr(out, C.c_double), x.size)
    if rc:
        raise ValueError(op)
    return out
