"""The reference-named entry points driven by a Fortran main loop in the
shape of main.f90:743-1149 (raishinsource_amd/fortran/raishin_main.f90),
against the reference's own trajectory and against the resident-model API."""
import os
import subprocess

import numpy as np
import pytest

from raishinsource_amd import configs, native

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def bitdiff(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    both_nan = np.isnan(a) & np.isnan(b)
    return int(np.count_nonzero((a.view(np.uint64) != b.view(np.uint64)) & ~both_nan))


def run_fortran(p, nsteps, out):
    if not os.path.exists(native.FORTRAN_MAIN):
        raise native.RaishinError(f"{native.FORTRAN_MAIN} is not built; run raishinsource_amd/_build.py")
    cmd = [native.FORTRAN_MAIN, str(p.imax), str(p.jmax), str(p.kmax), str(p.irec), str(p.irkt), repr(p.cfl),
           str(nsteps), str(out)]
    print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=240)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    raw = open(out, "rb").read()
    hdr = np.frombuffer(raw[:16], dtype=np.int32)
    ni, nj, nk, ns = (int(x) for x in hdr)
    off = 16
    dts = np.frombuffer(raw[off:off + 8 * ns], dtype=np.float64); off += 8 * ns
    time = float(np.frombuffer(raw[off:off + 8], dtype=np.float64)[0]); off += 8
    n = 9 * ni * nj * nk
    uri = np.frombuffer(raw[off:off + 8 * n], dtype=np.float64).reshape((9, ni, nj, nk), order="F"); off += 8 * n
    uu = np.frombuffer(raw[off:off + 8 * n], dtype=np.float64).reshape((9, ni, nj, nk), order="F"); off += 8 * n
    cnt = np.frombuffer(raw[off:off + 88], dtype=np.int64)
    return dts, time, uri, uu, cnt


def run_model(p, nsteps):
    g = native.setup_grid(p)
    m = native.setup_metric(p, g)
    uri = native.setup_torus(p, g, m, seed=0, noise=0.02)
    mod = native.Model(p, g, m)
    mod.set_state(uri)
    dts = []
    for _ in range(nsteps):
        t, dt = mod.step(1)
        dts.append(dt)
    w, u, t = mod.get_state()
    cnt = mod.counters()
    mod.close()
    return np.array(dts), t, w, u, cnt


def test_fortran_loop_matches_reference_trajectory(tmp_path):
    """10 steps of the 64x64 torus: the Fortran loop through the
    reference-named routines against the reference Fortran's trajectory."""
    p = configs.T2D.resolved()
    gold = dict(np.load(os.path.join(GOLD, "traj_t2d.npz")))
    dts, time, uri, uu, _ = run_fortran(p, len(gold["dt"]), tmp_path / "t2d.bin")
    assert bitdiff(dts, gold["dt"]) == 0
    j = p.nm1
    for jj in range(p.jmax):   # the axisymmetric run keeps every phi slice equal
        assert bitdiff(uri[:, :, jj, :], uri[:, :, j, :]) == 0
    for v in range(9):
        ref = gold["uri"][v]
        err = np.abs(uri[v, :, j, :] - ref).max() / max(np.abs(ref).max(), 1e-300)
        assert err <= 1e-11, (v, err)


@pytest.mark.parametrize("name,nsteps", [("t2d", 5), ("uppm", 6), ("ump5", 4), ("uweno", 3), ("umc", 4)])
def test_fortran_loop_equals_model_api(name, nsteps, tmp_path):
    """Same kernels, two boundaries: the routine-by-routine calls (dt and
    dt/dx computed by the Fortran host) give the resident-model result bit
    for bit, NaNs and counters included."""
    p = configs.ALL[name].resolved()
    fd, ft, fw, fu, fc = run_fortran(p, nsteps, tmp_path / f"{name}.bin")
    md, mt, mw, mu, mc = run_model(p, nsteps)
    assert bitdiff(fd, md) == 0
    assert ft == mt
    assert bitdiff(fw, mw) == 0
    assert bitdiff(fu, mu) == 0
    assert list(fc) == [mc[k] for k in ("recov_fallback", "recov_fail", "recov_neg_rho", "recov_neg_p",
                                        "recov_vsq_gt1", "recov_det0", "recov_wg_carry", "atm_floor", "pmm_scale",
                                        "pmm_floor", "guard")]
