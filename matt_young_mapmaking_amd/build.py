"""Build the in-tree native artefacts.

  matt_young_mapmaking_amd/libp2m.so   HIP library (gfx950) exporting include/pks2map_gpu.h
  matt_young_mapmaking_amd/bin/pks2map, bin/make_maptable
                                        drop-in Fortran executables (amdflang) over libp2m.so
  oracle/build/libp2mo*.so             CPU oracle (test infrastructure; built by oracle/Makefile)

Run: python -m matt_young_mapmaking_amd.build [--no-fortran] [--no-oracle]
"""
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
FSRC = os.path.join(PKG, "fortran")
LIB = os.path.join(PKG, "libp2m.so")
BIN = os.path.join(PKG, "bin")

HIP_SOURCES = ["p2m_api.hip", "p2m_kernels.hip", "p2m_maptable.hip"]
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # bit-exact FP64 geometry: no contraction, no fast-math
    "-ffp-contract=off", "-fno-fast-math",
    # hardware FP64/FP32 global atomics (device memory only)
    "-munsafe-fp-atomics",
    # no packed-FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError(f"build step failed: {' '.join(cmd[:3])} ...")
    return r.stdout


def hipcc():
    for c in ("/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def build_lib():
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + HIP_FLAGS + ["-I", CSRC, "-I", os.path.join(ROOT, "include")] + \
          [os.path.join(CSRC, s) for s in HIP_SOURCES] + ["-o", tmp]
    _run(cmd)
    os.replace(tmp, LIB)
    return LIB


def flang():
    for c in ("/opt/rocm/bin/amdflang", shutil.which("amdflang"), shutil.which("flang")):
        if c and os.path.exists(c):
            return c
    return None


def build_fortran():
    fc = flang()
    if fc is None:
        return None
    os.makedirs(BIN, exist_ok=True)
    mod = os.path.join(PKG, "fortran", "_mod")
    os.makedirs(mod, exist_ok=True)
    common = [os.path.join(FSRC, f) for f in ("p2m_iface.f90", "p2m_args.f90")]
    out = []
    for prog in ("pks2map", "make_maptable"):
        exe = os.path.join(BIN, prog)
        cmd = [fc, "-O2", "-J", mod] + common + [os.path.join(FSRC, prog + ".f90"),
               "-L", PKG, "-lp2m", "-Wl,-rpath,$ORIGIN/..", "-o", exe + ".tmp"]
        _run(cmd)
        os.replace(exe + ".tmp", exe)
        out.append(exe)
    return out


def build_oracle():
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    build_lib()
    if "--no-fortran" not in argv and os.path.exists(os.path.join(FSRC, "pks2map.f90")):
        build_fortran()
    if "--no-oracle" not in argv:
        build_oracle()


if __name__ == "__main__":
    main()
