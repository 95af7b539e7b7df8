#!/usr/bin/env python3
"""bench.py -- halos painted per second at HEALPix nside 4096 on one MI355X.

Workload (default "C3"): 2e7 halos isotropic on the sky, comoving distance
uniform in volume for 0.05 < z < 1.2, dn/dM ~ M^-1.9 on [1e12, 10^15.5] Msun,
painted with the 148 GHz tSZ table onto a full-sky nside-4096 HEALPix map and
the 1024^2 10-degree flat-sky map (pks2map CLI
`pks2map cat out tab 0.05 4096 0 0 1024 10 1.2 0 1 1`).

A step = one p2m_paint_device call on the HBM-resident catalogue: decode/cull,
binning, tile painting, flat-sky painting, maps resident on the device.  The
headline `value` is the FP32 + compensated painter over exactly --steps steps
(after --warmup), bracketed by device synchronisations.

Reported beside it, each timed as the median of >= 10 calls after >= 3
warm-ups:
  * the same painter with the catalogue handed over in page-locked host
    memory (p2m_paint: PCIe copy overlapped with the decode), and the raw
    PCIe copy rate that bounds it;
  * the FP64 tiled painter, and the plain FP64 painter (one work item per
    halo, FP64 global atomics: the 20x denominator), device-resident and from
    page-locked host memory;
  * the error of the FP32 and FP64 maps against the FP64 oracle painting the
    same catalogue on the host (per-pixel membership count and hash, max
    relative error, zero mismatches, integrated Y), and of the reference's
    own float32 arithmetic (the oracle's ref32 mode) against the same oracle;
  * cpu_baseline: the ref32 restatement (the reference's per-pixel work) on
    the host cores the run may use, with a single-core rate.

Prints one JSON line.
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from matt_young_mapmaking_amd import abi, catalog, devmem  # noqa: E402

METRIC = "halos painted/sec at nside 4096 (FP32+compensated vs FP64 GPU); max pixel rel err"
HBM_PEAK_GBS = 8000.0
CPU_SHARE = 16          # host cores one GPU's share of the box may use


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def measured_traffic(workload, n, tile_w, tile_wg):
    """HBM bytes per launch of the FP32 tile painter from the committed rocprofv3 --pmc
    summary (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration), when it
    covers this workload and tile shape; else None."""
    path = os.path.join(ROOT, "profiles", "k_tile_fast_C3_hbm_traffic.md")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    key, val = line.split(":", 1)[1].split("=")
                    wl, nn, tw, twg = key.strip().split(",")
                    if (wl, int(nn), int(tw), int(twg)) == (workload, n, tile_w, tile_wg):
                        return float(val)
    except (OSError, ValueError):
        return None
    return None


def new_painter(precision, nside, npix, hi, hf, tab, tile=None, membership=False):
    p = abi.Painter(nside=nside, npix_fs=npix, fov_deg=10.0, precision=precision, membership=membership)
    if tile and precision != abi.PREC_PLAIN_FP64:
        p.set_tile(*tile)
    p.set_run(zmin=0.05, zmax=1.2, chihview=0.0, mmin=1e13)
    p.load_table(0, hi, hf, tab)
    return p


def timed_steps(p, call, steps, warmup):
    """warmup untimed calls, then `steps` calls bracketed by device syncs:
    (seconds per step, per-step stats)."""
    for _ in range(warmup):
        p.clear()
        call()
    devmem.synchronize()
    stats = []
    t0 = time.perf_counter()
    for _ in range(steps):
        p.clear()
        call()
        stats.append(p.stats())
    devmem.synchronize()
    return (time.perf_counter() - t0) / steps, stats


def median_calls(p, call, reps, warmup):
    """Median wall time (s) of single calls (each call returns after its work
    is done: the library reads the paint counters back at the end)."""
    for _ in range(warmup):
        p.clear()
        call()
    devmem.synchronize()
    ts, stats = [], []
    for _ in range(reps):
        p.clear()
        devmem.synchronize()
        t = time.perf_counter()
        call()
        ts.append(time.perf_counter() - t)
        stats.append(p.stats())
    return float(np.median(ts)), stats


def errors(g, o, gmem=None, omem=None):
    """max per-pixel relative error over oracle-nonzero pixels, pixels zero in
    exactly one map, integrated-Y relative error, membership equality"""
    o = o.astype(np.float64, copy=False)
    nz = o != 0
    rel = np.abs(g[nz].astype(np.float64) - o[nz]) / np.abs(o[nz])
    out = {"max_rel_err": float(rel.max()) if rel.size else 0.0,
           "pixels_rel_err_gt_1e-6": int(np.count_nonzero(rel > 1e-6)),
           "zero_mismatch_pixels": int(np.count_nonzero((g != 0) != nz)),
           "integrated_y_rel_err": float(abs(g.astype(np.float64).sum() - o.sum()) / abs(o.sum()))}
    if gmem is not None and omem is not None:
        out["membership_pixels_differing"] = int(np.count_nonzero((gmem[0] != omem[0]) | (gmem[1] != omem[1])))
    return out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="C3")
    ap.add_argument("--n-halos", type=int, default=None)
    ap.add_argument("--reps", type=int, default=10, help="calls per median of the reported side measurements")
    ap.add_argument("--cpu-single-sample", type=int, default=1_000_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-oracle", action="store_true")
    ap.add_argument("--tile", default="1024,256", help="W,WG tile shape, or 'auto' for the library's autotuner "
                    "(default: the shape the committed profiles cover)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU sky-band "
              "sharding is a design, not a build", file=sys.stderr)
        sys.exit(2)
    reps, warm = max(10, args.reps), max(3, args.warmup)

    cfg = catalog.CONFIGS[args.workload]
    n = args.n_halos or cfg["n"]
    nside, npix = cfg["nside"], cfg["npix"]
    log(f"workload {args.workload}: {n} halos, nside {nside}, flat {npix}^2")
    t = time.perf_counter()
    halos = catalog.config_catalog(args.workload, n=n)
    log(f"catalogue generated in {time.perf_counter() - t:.1f}s")
    t = time.perf_counter()
    hi, hf, tab = abi.make_maptable(model=1, nu=148.0, alpha=0.0)
    log(f"148 GHz table made on the GPU in {time.perf_counter() - t:.2f}s")
    dev = devmem.DeviceArray(halos)
    log("catalogue resident in HBM")
    tile = tuple(int(x) for x in args.tile.split(",")) if args.tile and args.tile != "auto" else None

    # ---- headline: FP32 + compensated painter, HBM-resident catalogue
    p32 = new_painter(abi.PREC_FP32_COMP, nside, npix, hi, hf, tab, tile)
    sec32, st32 = timed_steps(p32, lambda: p32.paint_device(dev.address, n), args.steps, args.warmup)
    s = st32[-1]
    tile_ms = float(np.mean([x["ms_tile"] for x in st32]))
    log(f"fp32 (HBM-resident): {sec32 * 1e3:.3f} ms/step, stats {s}")

    # ---- the same painter from page-locked host memory (PCIe copy overlapped with the decode)
    pinned = p32.host_array(n)
    pinned[:] = halos
    med_pin, st_pin = median_calls(p32, lambda: p32.paint(pinned), reps, warm)
    h2d_ms = devmem.h2d_copy_ms(dev.address, pinned.ctypes.data, halos.nbytes)
    log(f"fp32 from pinned host memory: median {med_pin * 1e3:.3f} ms; raw pinned copy of {halos.nbytes} B "
        f"{h2d_ms:.3f} ms ({halos.nbytes / h2d_ms / 1e6:.1f} GB/s); in-call copy {st_pin[-1]['ms_h2d']:.3f} ms")
    med_hbm32, _ = median_calls(p32, lambda: p32.paint_device(dev.address, n), reps, 1)
    p32.close()
    del pinned

    # ---- FP64 tiled painter
    p64 = new_painter(abi.PREC_FP64, nside, npix, hi, hf, tab, tile)
    med64, _ = median_calls(p64, lambda: p64.paint_device(dev.address, n), reps, warm)
    hp64, _ = p64.maps(dtype=np.float64)
    p64.close()
    log(f"fp64 tiled: median {med64 * 1e3:.3f} ms")

    # ---- plain FP64 painter (the 20x denominator), HBM-resident and from pinned host memory
    plain = None
    if not args.no_plain:
        pp = new_painter(abi.PREC_PLAIN_FP64, nside, npix, hi, hf, tab)
        medp, stp = median_calls(pp, lambda: pp.paint_device(dev.address, n), reps, warm)
        ppin = pp.host_array(n)
        ppin[:] = halos
        medp_pin, _ = median_calls(pp, lambda: pp.paint(ppin), reps, warm)
        pp.close()
        del ppin
        plain = {"halos_per_s": n / medp, "ms_median": medp * 1e3, "calls": reps, "warmup": warm,
                 "ms_paint": float(np.median([x["ms_paint"] for x in stp])),
                 "pinned_host_halos_per_s": n / medp_pin, "pinned_host_ms_median": medp_pin * 1e3}
        log(f"plain fp64: median {medp * 1e3:.1f} ms (HBM), {medp_pin * 1e3:.1f} ms (pinned host)")

    # ---- errors against the FP64 oracle painting the same catalogue on the host
    err = None
    if not args.no_oracle:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import p2mo  # the oracle, as the checker
        nth = min(CPU_SHARE, len(os.sched_getaffinity(0)))
        pm = new_painter(abi.PREC_FP32_COMP, nside, npix, hi, hf, tab, tile, membership=True)
        pm.paint_device(dev.address, n)
        hp32, _ = pm.maps(dtype=np.float32)
        c32, h32, _, _ = pm.membership()
        pm.close()
        o = p2mo.Oracle(hi, hf, tab)
        t = time.perf_counter()
        ref = o.paint(halos, nside=nside, npix=npix, nthreads=nth, maps=True, membership=True)
        log(f"oracle (FP64, {nth} threads) painted the whole catalogue in {time.perf_counter() - t:.2f}s")
        omem = (ref["hp_count"], ref["hp_hash"])
        err = {"fp32": errors(hp32, ref["hp"], (c32, h32), omem),
               "fp64_tiled": errors(hp64, ref["hp"]),
               "oracle_n_upd_hp": int(ref["stats"]["n_upd_hp"])}
        del hp32, c32, h32, hp64
        t = time.perf_counter()
        r32 = o.paint_ref32(halos, nside=nside, npix=npix, nthreads=nth, maps=True, membership=True)
        log(f"ref32 painted the whole catalogue in {time.perf_counter() - t:.2f}s")
        err["ref32_reference_arithmetic"] = errors(r32["hp"], ref["hp"], (r32["hp_count"], r32["hp_hash"]), omem)
        del r32, ref
        log(f"errors vs oracle: {err}")

    # ---- roofline of the dominant kernel (the FP32 tile painter): algorithmic bytes per launch =
    # one read + one write of an FP32 map element per accepted (halo, pixel) pair, plus the
    # FP32 output map written once; duration = HIP events around that launch on the painter's stream
    n_upd = int(s["n_upd_hp"])
    npix_hp = 12 * nside * nside
    alg_bytes = n_upd * 2 * 4 + npix_hp * 4
    achieved = alg_bytes / (tile_ms * 1e-3) / 1e9
    pass_bytes = 28 * n + alg_bytes
    traffic = measured_traffic(args.workload, n, int(s["tile_w"]), int(s["tile_wg"]))

    # ---- CPU baseline: the reference's float32 arithmetic (ref32 restatement) on the host cores
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import p2mo  # the oracle's ref32 mode: timed here as the CPU baseline only
        nth = min(CPU_SHARE, len(os.sched_getaffinity(0)))
        o = p2mo.Oracle(hi, hf, tab)
        t = time.perf_counter()
        o.paint_ref32(halos, nside=nside, npix=npix, nthreads=nth, maps=True)
        dt_all = time.perf_counter() - t
        ns1 = min(args.cpu_single_sample, n)
        t = time.perf_counter()
        o.paint_ref32(halos[:ns1], nside=nside, npix=npix, nthreads=1, maps=True)
        dt_one = time.perf_counter() - t
        cpu = {"value": n / dt_all, "unit": "halos/s", "cores": nth, "kind": "port",
               "sample": f"the whole {args.workload} catalogue ({n} halos, same nside {nside} + {npix}^2 flat maps) "
                         f"through the ref32 restatement of the reference (float32 arithmetic, contiguous halo "
                         f"blocks per thread, thread-private maps summed at the end) on {nth} threads: "
                         f"{dt_all:.2f}s wall; single core: first {ns1} halos in {dt_one:.2f}s",
               "single_core_halos_per_s": ns1 / dt_one, "cpu_model": cpu_model(),
               "os_cpu_count": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0))}
        log(f"cpu ref32: {n / dt_all:.3e} halos/s on {nth} threads, {ns1 / dt_one:.3e} halos/s on one")

    value = n / sec32
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "halos/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec32 * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: {n} halos full sky, nside {nside} HEALPix + {npix}^2 flat-sky, "
                               f"148 GHz tSZ table, masses [{cfg['mlo']:.0e},{cfg['mhi']:.2e}] Msun, 0.05<z<1.2",
                   "n_halos": n, "nside": nside, "npix_fs": npix, "n_maps": 1,
                   "input": "catalogue resident in HBM when the timed region starts"},
        "max_rel_err": err["fp32"]["max_rel_err"] if err else None,
        "zero_mismatch_pixels": err["fp32"]["zero_mismatch_pixels"] if err else None,
        "integrated_y_rel_err": err["fp32"]["integrated_y_rel_err"] if err else None,
        "error_reference": "FP64 oracle painting the same catalogue on the host" if err else None,
        "errors_vs_oracle": err,
        "n_painted": int(s["n_painted"]),
        "n_upd_hp": n_upd,
        "n_fallback_fp64": int(s["n_fallback"]),
        "fp32_hbm_median_ms": med_hbm32 * 1e3,
        "pinned_host": {"halos_per_s": n / med_pin, "ms_median": med_pin * 1e3, "calls": reps, "warmup": warm,
                        "ms_h2d_in_call": float(np.median([x["ms_h2d"] for x in st_pin])),
                        "raw_pinned_copy_ms": h2d_ms, "raw_pinned_copy_GBs": halos.nbytes / h2d_ms / 1e6,
                        "bytes": int(halos.nbytes)},
        "fp64_tiled_halos_per_s": n / med64,
        "fp64_tiled_ms_median": med64 * 1e3,
        "plain_fp64": plain,
        "speedup_vs_plain_fp64": (plain["ms_median"] / (med_hbm32 * 1e3)) if plain else None,
        "speedup_vs_plain_fp64_pinned_host": (plain["pinned_host_ms_median"] / (med_pin * 1e3)) if plain else None,
        "precision_note": "per-pixel profile in f32, summed in f64 LDS tiles (compensated accumulation), "
                          "one f32 rounding per map pixel; centre angles, azimuth offsets, redshift inversion "
                          "and the membership edge band in f64",
        "phase_ms": {k: float(np.mean([x[k] for x in st32]))
                     for k in ("ms_decode", "ms_bin", "ms_paint", "ms_tile", "ms_finish")},
        "pass_alg_GBs": pass_bytes / sec32 / 1e9,
        "tile": {"w": int(s["tile_w"]), "wg": int(s["tile_wg"])},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_tile_fast<1>", "alg_bytes_per_launch": alg_bytes,
                     "kernel_ms": tile_ms},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
