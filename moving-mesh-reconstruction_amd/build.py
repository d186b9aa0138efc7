"""Builds libmmh_amd.so (HIP, gfx950) and the Fortran drop-in layer in-tree.

    python moving-mesh-reconstruction_amd/build.py [--no-fortran]

hipcc flags: -ffp-contract=off keeps every statement's rounding sequence (no fused multiply-add), so
FP64/FP32 results are bit-identical to the CPU restatement; -packed-fp32-ops is disabled for all
device code.
"""
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
FSRC = os.path.join(PKG, "fortran")
LIB = os.path.join(PKG, "libmmh_amd.so")
FLIB = os.path.join(PKG, "libmmh_fortran.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def run(cmd, **kw):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, **kw)


def build_hip():
    srcs = [os.path.join(CSRC, "mmh.hip")]
    tmp = LIB + ".tmp"
    run([HIPCC] + HIP_FLAGS + ["-I", os.path.join(REPO, "include")] + srcs + ["-o", tmp])
    os.replace(tmp, LIB)


def build_fortran():
    fc = shutil.which(os.environ.get("FC", "flang"))
    if not fc:
        print("flang not found: Fortran drop-in layer not built", flush=True)
        return False
    srcs = [os.path.join(FSRC, f) for f in ("mmh_gpu.f90", "mmh_legacy.f90")]
    if not all(os.path.exists(s) for s in srcs):
        return False
    bdir = os.path.join(PKG, "_build")
    os.makedirs(bdir, exist_ok=True)
    objs = []
    for s in srcs:
        o = os.path.join(bdir, os.path.basename(s).replace(".f90", ".o"))
        run([fc, "-O2", "-fPIC", "-c", s, "-o", o, "-module-dir", bdir])
        objs.append(o)
    tmp = FLIB + ".tmp"
    run([fc, "-shared", "-fPIC"] + objs + ["-o", tmp, "-L" + PKG, "-lmmh_amd", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, FLIB)
    # the reconstruction driver program (links the module layer and the HIP library)
    main_src = os.path.join(FSRC, "mmh_main.f90")
    if os.path.exists(main_src):
        exe = os.path.join(PKG, "mmh_main")
        mo = os.path.join(bdir, "mmh_main.o")
        run([fc, "-O2", "-c", main_src, "-o", mo, "-module-dir", bdir])
        run([fc, mo, "-o", exe + ".tmp", "-L" + PKG, "-lmmh_fortran", "-lmmh_amd", "-Wl,-rpath,$ORIGIN"])
        os.replace(exe + ".tmp", exe)
    return True


def main(argv):
    build_hip()
    if "--no-fortran" not in argv:
        build_fortran()


if __name__ == "__main__":
    main(sys.argv[1:])
