#!/usr/bin/env python3
"""Summaries of rocprofv3 output for the painter step, as markdown tables.

  trace   <dir>                       per-kernel mean duration (--kernel-trace)
  traffic <fetch_dir> <write_dir>     FETCH_SIZE / WRITE_SIZE per dispatch,
                                      scaled by the copy kernel's reading
  counters <dir> [<dir> ...]          SQ/TCC counters per kernel and the rates
                                      derived from them

Each <dir> is a `rocprofv3 ... --output-format csv -d <dir>` output folder of
profile_workload.py.  Counter values are those of a kernel's last dispatch.

Derived rates (MI355X: 256 CUs, 4 SIMDs each, 8 XCDs; GRBM_GUI_ACTIVE is
summed over the 8 XCDs, so one XCD's busy cycles are GRBM_GUI_ACTIVE / 8):
  VALU busy      = SQ_ACTIVE_INST_VALU * 4 / (4 * 256) / (GRBM_GUI_ACTIVE / 8)
  VALU lanes     = SQ_THREAD_CYCLES_VALU / (64 * SQ_ACTIVE_INST_VALU)
  LDS bank conflict = SQ_LDS_BANK_CONFLICT / SQ_LDS_IDX_ACTIVE
  waves resident = SQ_WAVE_CYCLES * 4 / (GRBM_GUI_ACTIVE / 8) / 256 per CU
                   (SQ_WAVE_CYCLES counts in units of 4 cycles)
"""
import collections
import csv
import glob
import sys

N_CU = 256


def _rows(d, kind):
    out = []
    for f in sorted(glob.glob(f"{d}/**/*{kind}.csv", recursive=True)):
        with open(f) as fh:
            out.extend(csv.DictReader(fh))
    if not out:
        sys.exit(f"no *{kind}.csv under {d}")
    return out


def _name(r):
    return r["Kernel_Name"].split("(")[0].replace("void ", "")


def trace(d):
    acc = collections.defaultdict(list)
    for r in _rows(d, "kernel_trace"):
        acc[_name(r)].append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-6)
    print("| Kernel | dispatches | mean ms | total ms |")
    print("|---|---|---|---|")
    for k, v in sorted(acc.items(), key=lambda kv: -sum(kv[1])):
        print(f"| `{k}` | {len(v)} | {sum(v) / len(v):.4f} | {sum(v):.3f} |")


def _counters(dirs):
    c = collections.defaultdict(dict)
    meta = {}
    for d in dirs:
        for r in _rows(d, "counter_collection"):
            k = _name(r)
            c[k][r["Counter_Name"]] = float(r["Counter_Value"])
            meta[k] = (r.get("VGPR_Count", ""), r.get("LDS_Block_Size", ""), r.get("Grid_Size", ""),
                       r.get("Workgroup_Size", ""))
    return c, meta


def traffic(fetch_dir, write_dir):
    c, _ = _counters([fetch_dir, write_dir])
    copy = c.get("k_copy_calib")
    if not copy:
        sys.exit("the copy kernel k_copy_calib is missing")
    # k_copy_calib moves `copy_mib` MiB each way per dispatch (1024 by default)
    true_kib = 1024 * 1024
    ff = copy["FETCH_SIZE"] / true_kib
    fw = copy["WRITE_SIZE"] / true_kib
    print(f"copy kernel: FETCH_SIZE {copy['FETCH_SIZE']:.1f} KiB, WRITE_SIZE {copy['WRITE_SIZE']:.1f} KiB "
          f"for {true_kib} KiB each way: factors {ff:.6f}, {fw:.6f}\n")
    print("| Kernel | FETCH raw KiB | WRITE raw KiB | fetch GB (scaled) | write GB (scaled) |")
    print("|---|---|---|---|---|")
    for k, v in c.items():
        if k == "k_copy_calib" or "FETCH_SIZE" not in v or "WRITE_SIZE" not in v:
            continue
        fg = v["FETCH_SIZE"] / ff * 1024 / 1e9
        wg = v["WRITE_SIZE"] / fw * 1024 / 1e9
        print(f"| `{k}` | {v['FETCH_SIZE']:.1f} | {v['WRITE_SIZE']:.1f} | {fg:.4f} | {wg:.4f} |")


def counters(dirs):
    c, meta = _counters(dirs)
    cols = ["VALU busy", "VALU lanes", "VALU insts/wave", "LDS insts/wave", "LDS bank conflict",
            "LDS atomics", "waves/CU", "wait any / wave cycles"]
    print("| Kernel | VGPR, LDS B | " + " | ".join(cols) + " |")
    print("|---|---|" + "---|" * len(cols))
    for k, v in c.items():
        g = v.get("GRBM_GUI_ACTIVE")
        waves = v.get("SQ_WAVES")
        if not g or not waves:
            continue
        xcd = g / 8.0
        f = {}
        if "SQ_ACTIVE_INST_VALU" in v:
            f["VALU busy"] = f"{100 * v['SQ_ACTIVE_INST_VALU'] * 4 / (4 * N_CU) / xcd:.1f}%"
            if "SQ_THREAD_CYCLES_VALU" in v:
                f["VALU lanes"] = f"{100 * v['SQ_THREAD_CYCLES_VALU'] / (64 * v['SQ_ACTIVE_INST_VALU']):.1f}%"
        if "SQ_INSTS_VALU" in v:
            f["VALU insts/wave"] = f"{v['SQ_INSTS_VALU'] / waves:.0f}"
        if "SQ_INSTS_LDS" in v:
            f["LDS insts/wave"] = f"{v['SQ_INSTS_LDS'] / waves:.0f}"
        if "SQ_LDS_BANK_CONFLICT" in v and v.get("SQ_LDS_IDX_ACTIVE"):
            f["LDS bank conflict"] = f"{100 * v['SQ_LDS_BANK_CONFLICT'] / v['SQ_LDS_IDX_ACTIVE']:.1f}%"
        if "SQ_INSTS_LDS_ATOMIC" in v:
            f["LDS atomics"] = f"{v['SQ_INSTS_LDS_ATOMIC']:.3g}"
        if "SQ_WAVE_CYCLES" in v:
            f["waves/CU"] = f"{v['SQ_WAVE_CYCLES'] * 4 / xcd / N_CU:.1f}"
            if "SQ_WAIT_ANY" in v:
                f["wait any / wave cycles"] = f"{100 * v['SQ_WAIT_ANY'] / v['SQ_WAVE_CYCLES']:.0f}%"
        m = meta[k]
        print(f"| `{k}` | {m[0]}, {m[1]} | " + " | ".join(f.get(x, "") for x in cols) + " |")
    print()
    print("Raw values (last dispatch):\n")
    names = sorted({n for v in c.values() for n in v})
    for k, v in c.items():
        print(f"- `{k}`: " + ", ".join(f"{n} {v[n]:.0f}" for n in names if n in v))


def main():
    if len(sys.argv) < 3:
        sys.exit(__doc__)
    cmd, args = sys.argv[1], sys.argv[2:]
    if cmd == "trace":
        trace(args[0])
    elif cmd == "traffic":
        traffic(args[0], args[1])
    elif cmd == "counters":
        counters(args)
    else:
        sys.exit(__doc__)


if __name__ == "__main__":
    main()
