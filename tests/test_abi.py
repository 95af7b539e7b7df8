"""The C-ABI library without a GPU: it loads, exports every function the
public header declares, validates arguments before touching a device, and
reports a missing device as an error (never a CPU fallback)."""
import ctypes as C
import os
import re
import subprocess

import pytest

from matt_young_mapmaking_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pks2map_gpu.h")


def declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(p2m_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = abi.load()
    names = declared_functions()
    assert len(names) >= 18
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(abi.EXPORTS)
    nm = subprocess.run(["nm", "-D", "--defined-only", abi.LIB_PATH], capture_output=True, text=True).stdout
    for n in names:
        assert re.search(rf"\bT {n}\b", nm), n


def test_library_does_not_link_the_oracle():
    out = subprocess.run(["ldd", abi.LIB_PATH], capture_output=True, text=True).stdout
    assert "p2mo" not in out
    nm = subprocess.run(["nm", "-D", abi.LIB_PATH], capture_output=True, text=True).stdout
    assert "p2mo_" not in nm


def test_argument_validation_before_device():
    bad = [dict(nside=3), dict(nside=16384), dict(n_maps=0), dict(n_maps=7), dict(model=4),
           dict(profile=3), dict(precision=5), dict(npix_fs=-1)]
    for kw in bad:
        with pytest.raises(abi.P2MError) as e:
            abi.Painter(**kw)
        assert e.value.code == -1, kw


def test_no_device_is_an_error():
    with pytest.raises(abi.P2MError) as e:
        abi.Painter(nside=64)
    assert e.value.code in (-8, -3)
    with pytest.raises(abi.P2MError):
        abi.make_maptable(nchit=3, nmht=3, nrt=3)
    with pytest.raises(abi.P2MError) as e:
        abi.selftest_copy(1 << 20)
    assert e.value.code in (-8, -3)
    for nb, reps in ((17, 1), (0, 1), (1 << 20, 0), ((1 << 34) + 16, 1)):
        with pytest.raises(abi.P2MError) as e:
            abi.selftest_copy(nb, reps)
        assert e.value.code == -1


def test_strerror_and_version():
    L = abi.load()
    for code, word in ((0, b"ok"), (-1, b"invalid"), (-2, b"size"), (-4, b"alloc"), (-8, b"GPU")):
        assert word in L.p2m_strerror(code)
    assert b"gfx950" in L.p2m_version()


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(abi, "_lib", None)
    monkeypatch.setattr(abi, "LIB_PATH", str(tmp_path / "nope.so"))
    with pytest.raises(RuntimeError, match="not built"):
        abi.load()
