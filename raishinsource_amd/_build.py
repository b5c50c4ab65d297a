"""Build libraishin_hip.so in-tree for gfx950.

Kernels and the C-ABI are compiled by hipcc, the host set-up by g++.  Both
keep IEEE FP64 semantics that the parity with the reference relies on:
-ffp-contract=off (no FMA contraction), no fast-math.  Packed FP32 ops are
disabled for all device code.
"""
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
             "-Wno-unused-command-line-argument"]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math"]

UNITS = [("rh_kernels.hip", "hip"), ("rh_capi.cpp", "hip"), ("rh_setup.cpp", "host")]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError("build failed: " + " ".join(cmd))
    return r.stdout


def build(verbose=False):
    os.makedirs(OUT, exist_ok=True)
    objs = []
    for src, kind in UNITS:
        s = os.path.join(CSRC, src)
        o = os.path.join(OUT, src + ".o")
        if kind == "hip":
            cmd = [HIPCC] + HIP_FLAGS + ["-c", s, "-o", o]
        else:
            cmd = ["g++"] + HOST_FLAGS + ["-c", s, "-o", o]
        out = _run(cmd)
        if verbose and out:
            print(out)
        objs.append(o)
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", LIB] + objs)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
