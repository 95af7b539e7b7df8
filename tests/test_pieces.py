"""The ring tiles (ring_tiles, tile_span in p2m_dev.h) stay inside the map
for every nside 1..8192 and tile width 64..2048.  The enumeration of a ring
run's tile pieces (p2m_pieces.h), which the
binning uses to cut every (halo, ring) run at the tile boundaries: each piece
lies in one tile, the pieces follow the run from its first pixel in ring
order, wrapping past the ring's end once, and cover it exactly; the emit
pass writes every run's first piece with first_two_pieces, which must give
run_pieces' first pieces.  Checked on
the host over every run start and length (strided on long rings), ring
lengths that are and are not multiples of the tile width, and tile widths
64..2048."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "matt_young_mapmaking_amd", "csrc")


def test_run_pieces_partition_the_run(tmp_path):
    hipcc = "/opt/rocm/bin/hipcc" if os.path.exists("/opt/rocm/bin/hipcc") else shutil.which("hipcc")
    if not hipcc:
        pytest.skip("hipcc not found")
    exe = tmp_path / "pieces_check"
    subprocess.run([hipcc, "-O2", "-std=c++17", "-x", "hip", "--offload-arch=gfx950", "-I", CSRC,
                    os.path.join(ROOT, "tests", "pieces_check.cpp"), "-o", str(exe)],
                   check=True, capture_output=True, timeout=300)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout
    assert "0 mismatches" in r.stdout


def test_fast_ring_runs_equal_fp64_runs(tmp_path):
    """disc_ring_run_fast (FP32 arc half-width, FP64 re-decide band) returns
    the FP64 run of pixels_on_edge for random discs of radius 1e-6..2 rad at
    nside 1..8192, polar centres included (the binning uses it for every
    (halo, ring) item)."""
    hipcc = "/opt/rocm/bin/hipcc" if os.path.exists("/opt/rocm/bin/hipcc") else shutil.which("hipcc")
    if not hipcc:
        pytest.skip("hipcc not found")
    exe = tmp_path / "runs_check"
    subprocess.run([hipcc, "-O2", "-std=c++17", "-x", "hip", "--offload-arch=gfx950", "-ffp-contract=off",
                    "-I", CSRC, os.path.join(ROOT, "tests", "runs_check.cpp"), "-o", str(exe)],
                   check=True, capture_output=True, timeout=300)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout
    assert " 0 mismatches" in r.stdout


def test_ring_tiles_stay_inside_the_map(tmp_path):
    """Every tile of every ring, for nside 1..8192 and W = 64..2048: its
    first map pixel and width keep it inside its ring and the map, and the
    tiles of a ring cover it exactly (the bounds the tile painter's stores
    rely on)."""
    hipcc = "/opt/rocm/bin/hipcc" if os.path.exists("/opt/rocm/bin/hipcc") else shutil.which("hipcc")
    if not hipcc:
        pytest.skip("hipcc not found")
    exe = tmp_path / "tiles_check"
    subprocess.run([hipcc, "-O2", "-std=c++17", "-x", "hip", "--offload-arch=gfx950", "-I", CSRC,
                    os.path.join(ROOT, "tests", "tiles_check.cpp"), "-o", str(exe)],
                   check=True, capture_output=True, timeout=300)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout
    assert " 0 mismatches" in r.stdout
