#!/usr/bin/env python3
"""One painting workload for a profiler to wrap (rocprofv3 --kernel-trace or
--pmc): the copy-kernel calibration, then `--steps` FP32 painter steps on an
HBM-resident catalogue exactly as bench.py times them (no work skipped).

Prints one JSON line with what the profile summary needs: the copy kernel's
byte count, and per step the update count, map size and tile shape.

    python3 matt_young_mapmaking_amd/profile_workload.py --workload C3 --tile 512,256 --steps 2
"""
import argparse
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from matt_young_mapmaking_amd import abi, catalog, devmem  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="C3")
    ap.add_argument("--n-halos", type=int, default=None)
    ap.add_argument("--tile", default="512,256")
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--copy-mib", type=int, default=1024)
    args = ap.parse_args()
    cfg = catalog.CONFIGS[args.workload]
    n = args.n_halos or cfg["n"]
    copy_bytes = args.copy_mib << 20
    copy_ms = abi.selftest_copy(copy_bytes, reps=3)
    print(f"[profile] copy {copy_bytes} B: {copy_ms:.3f} ms, {2 * copy_bytes / copy_ms / 1e6:.1f} GB/s", flush=True)
    halos = catalog.config_catalog(args.workload, n=n)
    hi, hf, tab = abi.make_maptable(model=1, nu=148.0, alpha=0.0)
    dev = devmem.DeviceArray(halos)
    tw, twg = (int(x) for x in args.tile.split(","))
    steps = []
    with abi.Painter(nside=cfg["nside"], npix_fs=cfg["npix"], fov_deg=10.0, precision=abi.PREC_FP32_COMP) as p:
        p.set_tile(tw, twg)
        p.set_run(zmin=0.05, zmax=1.2, chihview=0.0, mmin=1e13)
        p.load_table(0, hi, hf, tab)
        for k in range(args.steps):
            p.clear()
            p.paint_device(dev.address, n)
            s = p.stats()
            steps.append({"ms_tile": s["ms_tile"], "ms_total": s["ms_total"], "n_upd_hp": int(s["n_upd_hp"]),
                          "n_segments": int(s["n_segments"]), "n_fallback": int(s["n_fallback"])})
            print(f"[profile] step {k}: {s['ms_total']:.3f} ms, tile kernel {s['ms_tile']:.3f} ms", flush=True)
    devmem.synchronize()
    print(json.dumps({"workload": args.workload, "n_halos": n, "nside": cfg["nside"], "npix_hp": 12 * cfg["nside"] ** 2,
                      "tile": [tw, twg], "copy_bytes": copy_bytes, "copy_ms": copy_ms, "steps": steps}))


if __name__ == "__main__":
    main()
