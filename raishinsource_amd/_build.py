"""Build libraishin_hip.so in-tree for gfx950.

Kernels and the C-ABI are compiled by hipcc, the host set-up by g++.  Both
keep IEEE FP64 semantics that the parity with the reference relies on:
-ffp-contract=off (no FMA contraction), no fast-math, IEEE FP64 divides and
square roots.  Packed FP32 ops are disabled for all device code; FP32
divides and square roots (the mixed form only) use the hardware's
approximations.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libraishin_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("RAISHIN_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
             "-fno-fast-math", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-command-line-argument",
             # keep MachineLICM from hoisting every FP64 constant and plane
             # offset out of the kernels' loops (the PPM sweep's walk along j
             # and k, the Newton iterations): held across the loop they
             # spill registers
             "-mllvm", "-disable-machine-licm",
             # float divides and square roots (the mixed form's sweeps only:
             # FP64 ones stay IEEE, and the FP64 kernels compile to the same
             # code) by the hardware reciprocal / square root, a few ulp
             "-fno-hip-fp32-correctly-rounded-divide-sqrt"]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-fopenmp"]

FC = os.environ.get("FC_FLANG", "/opt/rocm/lib/llvm/bin/flang")
FORTRAN = ["raishin_c.f90", "raishin_wrappers.f90", "raishin_main.f90"]
FORTRAN_MAIN = os.path.join(HERE, "fortran", "raishin_main")

UNITS = [("rh_kernels.hip", "hip"), ("rh_model.cpp", "hip"), ("rh_capi.cpp", "hip"), ("rh_ref.cpp", "hip"),
         ("rh_batch.cpp", "hip"), ("rh_noalloc.hip", "hip"), ("rh_torus.hip", "hip"),
         ("rh_setup.cpp", "host"), ("rh_io.cpp", "host")]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError("build failed: " + " ".join(cmd))
    return r.stdout


def build(verbose=False):
    os.makedirs(OUT, exist_ok=True)
    objs, cmds = [], []
    for src, kind in UNITS:
        s = os.path.join(CSRC, src)
        o = os.path.join(OUT, src + ".o")
        if kind == "hip":
            cmds.append([HIPCC] + HIP_FLAGS + ["-c", s, "-o", o])
        else:
            cmds.append(["g++"] + HOST_FLAGS + ["-c", s, "-o", o])
        objs.append(o)
    # the units are independent: compile them side by side
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(len(cmds), 4)) as ex:
        for out in ex.map(_run, cmds):
            if verbose and out:
                print(out)
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", LIB] + objs + ["-fopenmp"])
    build_fortran(verbose)
    return LIB


def build_fortran(verbose=False):
    """The iso_c_binding module, the reference-named wrappers and the Fortran
    main loop, linked against the library (found next to it at run time)."""
    fdir = os.path.join(HERE, "fortran")
    fobjs = []
    for src in FORTRAN:
        o = os.path.join(OUT, src + ".o")
        out = _run([FC, "-O2", "-J", OUT, "-c", os.path.join(fdir, src), "-o", o])
        if verbose and out:
            print(out)
        fobjs.append(o)
    _run([FC, "-o", FORTRAN_MAIN] + fobjs + ["-L" + HERE, "-lraishin_hip", "-Wl,-rpath,$ORIGIN/.."])
    return FORTRAN_MAIN


if __name__ == "__main__":
    print(build(verbose=True))
