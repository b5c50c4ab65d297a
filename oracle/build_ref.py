#!/usr/bin/env python3
"""Build the reference Fortran hot-path routines into ``oracle/_ref/<cfg>/libref.so``.

Test infrastructure only.  The sources are compiled where they lie under
/root/reference (nothing is copied into the repository); the one generated
file is the configuration module: the reference's own ``pram.f90`` with its
parameter values rewritten for the chosen run, which is how the reference
itself is configured (``prepare.py:102-137`` rewrites the same file).

Compiled (none of them include ``mpif.h``): pram, coord, grid, physfunc,
caluu, calflx, calcha, hll, ct, calsf, rkt, recov, correction, cdtcfl, func,
rec.  Not compiled: bnd.f90 (``bnd.f90:32`` includes ``mpif.h``, no MPI in
this image), main.f90, mdgrmhd.f90, mpisub.f90, restar.f90 (same reason).
Flags follow ``Makefile_xgrmhd:6`` (``-O2``) with ``-fPIC`` for a shared
object; no ``-march`` so x86-64 code has no FMA contraction.

Usage:  python oracle/build_ref.py c1 t2d ...   (names from raishinsource_amd/configs.py)
"""
import os
import re
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from raishinsource_amd.configs import ALL  # noqa: E402

REF = os.environ.get("RAISHIN_REFERENCE", "/root/reference")
FC = os.environ.get("FC_FLANG", "/opt/rocm/lib/llvm/bin/flang")
SOURCES = ["pram", "func", "physfunc", "coord", "grid", "caluu", "calflx",
           "calcha", "hll", "ct", "calsf", "rkt", "recov", "correction",
           "cdtcfl", "rec"]


def fnum(x):
    """A real(8) literal that parses back to exactly ``x``."""
    s = repr(float(x))
    if "e" in s:
        mant, ex = s.split("e")
        return f"{mant}d{ex}"
    return s + "d0"


def configured_pram(p):
    p = p.resolved()
    with open(os.path.join(REF, "pram.f90")) as f:
        text = f.read()
    ints = dict(imax=p.imax, jmax=p.jmax, kmax=p.kmax, iprocs=1, jprocs=1,
                kprocs=1, icres=1, iter=p.iter, ix1=p.ix1, ix2=p.ix2,
                ix3=p.ix3, model=p.model, metric=p.metric, ieos=p.ieos,
                icha=p.icha, iwvec=p.iwvec, irkt=p.irkt, irec=p.irec,
                ihll=p.ihll, ict=p.ict)
    reals = dict(cfl=p.cfl, dmin=p.dmin, pmin=p.pmin, pmax=p.pmax,
                 gam=p.gam, akm=p.akm, xmin=p.xmin, xmax=p.xmax,
                 ymin=p.ymin, ymax=p.ymax, zmin=p.zmin, zmax=p.zmax)
    for name, val in ints.items():
        text, n = re.subn(rf"(\b{name}\s*=\s*)([^,!\n]+)",
                          lambda m: m.group(1) + str(val), text, count=1)
        assert n == 1, name
    for name, val in reals.items():
        text, n = re.subn(rf"(::\s*{name}\s*=\s*)([^!\n]+)",
                          lambda m: m.group(1) + fnum(val), text, count=1)
        assert n == 1, name
    for name in ("iboux1in", "iboux1ot", "iboux2in", "iboux2ot",
                 "iboux3in", "iboux3ot"):
        vals = ", ".join(str(v) for v in getattr(p, name))
        text, n = re.subn(rf"({name}\(9\)\s*=\s*)\(/[^/]*/\)",
                          lambda m: m.group(1) + f"(/{vals}/)", text, count=1)
        assert n == 1, name
    return text


def _run(cmd, cwd):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError("failed: " + " ".join(cmd))


def build(name):
    p = ALL[name]
    out = os.path.join(HERE, "_ref", name)
    os.makedirs(out, exist_ok=True)
    with open(os.path.join(out, "pram.f90"), "w") as f:
        f.write(configured_pram(p))
    objs = []
    for src in SOURCES:
        path = os.path.join(out, "pram.f90") if src == "pram" else os.path.join(REF, src + ".f90")
        obj = os.path.join(out, src + ".o")
        _run([FC, "-O2", "-fPIC", "-J", out, "-c", path, "-o", obj], out)
        objs.append(obj)
    lib = os.path.join(out, "libref.so")
    _run([FC, "-shared", "-o", lib] + objs, out)
    for o in objs:
        os.remove(o)
    return lib


if __name__ == "__main__":
    if not os.path.isdir(REF):
        sys.exit(f"reference sources not found at {REF}")
    from concurrent.futures import ThreadPoolExecutor
    names = sys.argv[1:] or ["uppm", "ump5", "uweno", "t2d", "t3d", "c1"]
    with ThreadPoolExecutor(max_workers=min(8, len(names))) as ex:
        for lib in ex.map(build, names):
            print("built", lib)
