"""Build libshcfd.so (HIP, gfx950) in-tree and the Fortran host driver.

    python -m shared_cfd_amd.build            # library + Fortran driver
    python -m shared_cfd_amd.build --lib      # library only

The library is compiled with hipcc for gfx950 only, with FMA contraction
off (the kernels state their evaluation order explicitly) and packed FP32
instructions disabled.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(ROOT, "_build", "obj")
LIB = os.path.join(HERE, "libshcfd.so")
FORT = os.path.join(HERE, "fortran")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-fno-fast-math", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-variable", "-Wno-unused-but-set-variable",
]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l and "is not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    return out


# sources whose kernels take the species count as a template parameter: each
# is compiled once per instantiation group (SHCFD_INST_GROUP, see
# csrc/shcfd_kernels.h), so the 16 species counts compile in parallel
GROUPED = ("k_flow.hip", "k_implicit.hip", "k_sweep16.hip", "k_chem.hip")
INST_GROUPS = 5


def build_lib(jobs: int = 8, verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    srcs = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip") or f.endswith(".cpp"))
    units = []
    for src in srcs:
        if src in GROUPED:
            units += [(src, g) for g in range(INST_GROUPS)]
        else:
            units.append((src, None))
    # the largest units first, so the parallel build ends together
    units.sort(key=lambda u: (u[0] != "k_chem.hip", u[1] is None))

    def comp(unit):
        src, g = unit
        base = os.path.splitext(src)[0] + ("" if g is None else f"_g{g}")
        obj = os.path.join(OBJ, base + ".o")
        extra = [] if g is None else [f"-DSHCFD_INST_GROUP={g}"]
        cmd = [HIPCC, *HIP_FLAGS, *extra, "-c", os.path.join(CSRC, src), "-o", obj]
        out = _run(cmd)
        if verbose and out.strip():
            print(out)
        return obj

    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(comp, units))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


def build_fortran() -> str | None:
    """Fortran host module + drivers (flang), linked against libshcfd.so."""
    flang = shutil.which("flang") or "/opt/rocm/llvm/bin/flang"
    if not os.path.exists(flang) and not shutil.which("flang"):
        return None
    out_dir = os.path.join(ROOT, "_build", "fortran")
    os.makedirs(out_dir, exist_ok=True)
    mod = os.path.join(FORT, "shcfd_f.f90")
    drivers = [f for f in sorted(os.listdir(FORT)) if f.startswith("main_") and f.endswith(".f90")]
    # unformatted files big-endian, as the reference builds them (-convert big_endian,
    # store/architecture/PC/Makefile.head:4)
    fflags = ["-O2", "-fconvert=big-endian"]
    _run([flang, *fflags, "-c", mod, "-module-dir", out_dir, "-o", os.path.join(out_dir, "shcfd_f.o")])
    exe = None
    for drv in drivers:
        exe = os.path.join(HERE, drv[:-4])
        _run([flang, *fflags, "-I", out_dir, os.path.join(FORT, drv), os.path.join(out_dir, "shcfd_f.o"),
              "-L" + HERE, "-lshcfd", "-Wl,-rpath,$ORIGIN", "-o", exe])
    return exe


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib", action="store_true", help="library only")
    ap.add_argument("-j", type=int, default=8)
    ap.add_argument("-v", action="store_true")
    a = ap.parse_args(argv)
    print("built", build_lib(a.j, a.v))
    if not a.lib and os.path.isdir(FORT):
        exe = build_fortran()
        if exe:
            print("built", exe)


if __name__ == "__main__":
    sys.exit(main())
