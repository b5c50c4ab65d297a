#!/usr/bin/env python3
"""Write the golden vectors of tests/golden/ from the reference Fortran.

Runs in the container that holds /root/reference: builds the reference
routines (oracle/build_ref.py) and records, for small grids, the inputs and
outputs of the hot-path routines and of short step sequences driven in
main.f90's order.  The vectors are data only (npz).  Regenerate with
    python tests/golden/make_golden.py [unit] [c1]
"""
import hashlib
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, ROOT)
import ctypes as C  # noqa: E402

from harness import RefStepper, fzeros, grid_and_metric, have_ref, synthetic_state  # noqa: E402
from raishinsource_amd import configs, native  # noqa: E402

UNIT = ["gppm", "gmp5", "gweno"]  # 14x11x12 style grids, one per reconstruction


def unit_vectors(name):
    r, g, m = grid_and_metric(name)
    p = r.p
    ni, nj, nk = p.imax, p.jmax, p.kmax
    nm0, nm1, B = p.nm0, p.nm1, r.bounds()
    uri = synthetic_state(p, seed=7)
    out = {"uri": uri}
    uu = fzeros(9, ni, nj, nk)
    r.call("caluu1", uri, uu, m["gcov"], m["gcon"], m["detg"], g["x1"], *B)
    out["caluu1_uu"] = uu
    names = ["uriir", "urijr", "urikr", "uriil", "urijl", "urikl", "uuir", "uujr", "uukr", "uuil", "uujl", "uukl",
             "wwir", "wwjr", "wwkr", "wwil", "wwjl", "wwkl"]
    F = {n: fzeros(9, ni, nj, nk) for n in names}
    r.call("rec2", uri, g["x1"], g["x2"], g["x3"], g["x1a"], g["x2a"], g["x3a"], g["dx1"], g["dx2"], g["dx3"],
           g["dx1b"], g["dx2b"], g["dx3b"], g["x1b"], g["x2b"], g["x3b"], g["x1ab"], g["x2ab"], g["x3ab"],
           *[F[n] for n in names], m["gcovi"], m["gcovj"], m["gcovk"], m["gconi"], m["gconj"], m["gconk"],
           m["detgi"], m["detgj"], m["detgk"], nm0, *B)
    for n in ("uriir", "uriil", "urikr", "urikl", "wwir", "wwkl"):
        out["rec2_" + n] = F[n]
    c = [fzeros(ni, nj, nk) for _ in range(6)]
    r.call("calcha", *[F[n] for n in names[:6]], m["gcovi"], m["gcovj"], m["gcovk"], m["gconi"], m["gconj"],
           m["gconk"], *c, nm0, *B)
    out["calcha_c"] = np.stack(c)
    ww = fzeros(3, 9, ni, nj, nk)
    r.call("hll", ww, *[F[n] for n in names[6:]], *c, nm0, *B)
    out["hll_ww"] = ww.copy(order="F")
    r.call("ct", ww, ww.copy(order="F"), uri, nm0, nm1, *B)
    out["ct_ww"] = ww
    tenr = fzeros(4, 4, ni, nj, nk)
    sf = fzeros(4, ni, nj, nk)
    r.call("caltenr", uri, tenr, m["gcov"], m["gcon"], *B)
    r.call("calsf", sf, tenr, m["christ"], m["detg"], *B)
    out["calsf_sf"] = sf
    uh = uu.copy(order="F")
    uh[:, nm1:ni - nm1, nm1:nj - nm1, nm1:nk - nm1] *= (1 + 1e-3 * np.sin(np.arange(9)))[:, None, None, None]
    out["recov_uin"] = uh.copy(order="F")
    w2 = uri.copy(order="F")
    r.call("recov", uh, w2, m["gcov"], m["gcon"], m["detg"], g["x1"], g["x2"], g["x3"], nm1, *B)
    out["recov_uri"] = w2
    out["recov_uu"] = uh
    for fn in ("atmosphere1a", "pminmax2b"):
        w3 = uri.copy(order="F")
        r.call(fn, uu, w3, m["gcov"], m["gcon"], m["detg"], g["x1"], *B)
        out[fn + "_uri"] = w3
    d = [C.c_double(0.0) for _ in range(4)]
    r.call("cdtcfl4", uri, m["gcov"], m["gcon"], d[0], g["dx1b"], g["dx2b"], g["dx3b"], d[1], d[2], d[3], nm1, *B)
    out["cdtcfl4"] = np.array([x.value for x in d])
    return out


def trajectory(name, nsteps, seed=0, snap=()):
    """Torus from the package set-up, stepped by the reference in main.f90's order."""
    r, g, m = grid_and_metric(name)
    p = r.p
    uri = native.setup_torus(p, g, m, seed=seed)
    uu = fzeros(9, p.imax, p.jmax, p.kmax)
    r.call("caluu1", uri, uu, m["gcov"], m["gcon"], m["detg"], g["x1"], *r.bounds())
    out = {"uri0": uri.copy(order="F")}
    rs = RefStepper(name, g, m)
    t = 0.0
    dts = []
    for s in range(nsteps):
        dt, t = rs.step(uu, uri, t)
        dts.append(dt)
        if s + 1 in snap:
            out[f"uri_{s + 1}"] = uri.copy(order="F")
            out[f"uu_{s + 1}"] = uu.copy(order="F")
    out["uri"] = uri
    out["uu"] = uu
    out["dt"] = np.array(dts)
    out["time"] = t
    return out


def digest(a):
    """sha256 of an array's FP64 bit patterns (Fortran order)."""
    return hashlib.sha256(np.asfortranarray(a).tobytes(order="F")).hexdigest()


def c1_trajectory():
    """The reference's own 2D torus (128 x 128, PPM, RK2) stepped 100 times:
    digests of the full state after 10 steps and the interior phi slice of
    the primitives after 100."""
    tr = trajectory("c1", 100, snap=(10,))
    j = configs.C1.nm1
    return dict(uri0_digest=digest(tr["uri0"]), uri10_digest=digest(tr["uri_10"]), uu10_digest=digest(tr["uu_10"]),
                uri100=tr["uri"][:, :, j, :], dt=tr["dt"], time=np.array(tr["time"]))


def main(which):
    for n in UNIT + ["t2d", "c1"]:
        if not have_ref(n):
            subprocess.run([sys.executable, os.path.join(ROOT, "oracle", "build_ref.py"), n], check=True,
                           stdout=subprocess.DEVNULL)
    if "c1" in which:
        np.savez_compressed(os.path.join(HERE, "traj_c1.npz"), **c1_trajectory())
        print("wrote traj_c1")
    if "unit" not in which:
        return
    for n in UNIT:
        np.savez_compressed(os.path.join(HERE, f"unit_{n}.npz"), **unit_vectors(n))
        print("wrote", n)
    tr = trajectory("t2d", 10)
    # the 2D run is axisymmetric: keep the interior phi slice only
    j = configs.T2D.nm1  # 0-based index of the interior slice
    np.savez_compressed(os.path.join(HERE, "traj_t2d.npz"), uri0=tr["uri0"][:, :, j, :], uri=tr["uri"][:, :, j, :],
                        uu=tr["uu"][:, :, j, :], dt=tr["dt"])
    print("wrote traj_t2d")


if __name__ == "__main__":
    main(sys.argv[1:] or ["unit", "c1"])
