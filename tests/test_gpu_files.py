"""Checkpoint and restart of device-resident state in the reference's files
(restar.f90): a model stored, restarted into a fresh model and stepped on is
bit-identical to the model stepped straight through; the analysis file of a
model equals the host writer's; and the Fortran main loop's store1 /
restar2 / output1 (icres=1 restart, main.f90:560-571) resume its trajectory
bit for bit."""
import os
import subprocess

import numpy as np
import pytest

from raishinsource_amd import configs, native

pytestmark = pytest.mark.gpu


def bitdiff(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    both_nan = np.isnan(a) & np.isnan(b)
    return int(np.count_nonzero((a.view(np.uint64) != b.view(np.uint64)) & ~both_nan))


def fresh(p, mixed=False):
    p = p.resolved()
    g = native.setup_grid(p)
    m = native.setup_metric(p, g)
    mod = native.Model(p, g, m)
    mod.set_precision(mixed)
    return mod, g, m


# NaN-free trajectories only: where both wave speeds of an interface are NaN
# (the MP5 torus), hll.f90 keeps the flux slot of the previous step, state
# that the reference's restart file does not hold either; a restarted run
# then starts, like the reference's (main.f90:372-381), from zeroed fluxes.
@pytest.mark.parametrize("name,mixed", [("t2d", False), ("uppm", True), ("umc", False)])
def test_model_checkpoint_resumes_bit_for_bit(name, mixed, tmp_path):
    p = configs.ALL[name].resolved()
    a, g, m = fresh(p, mixed)
    a.set_state(native.setup_torus(p, g, m, seed=1, noise=0.02))
    a.step(3)
    w3, u3, t3 = a.get_state()
    path = tmp_path / "restart000.outdat"
    a.store1(path, it0=4, ih0=3, nd=1)
    a.step(4)
    wa, ua, ta = a.get_state()
    assert not np.isnan(wa).any()
    b, _, _ = fresh(p, mixed)
    u0, uri0, it0, ih0, nd = b.restar2(path)
    assert (it0, ih0, nd) == (4, 3, 1)
    assert bitdiff(u0, u3) == 0 and bitdiff(uri0, w3) == 0  # NULL u0/uri0: the current state stands in
    wb, ub, tb = b.get_state()
    assert tb == t3 and bitdiff(wb, w3) == 0 and bitdiff(ub, u3) == 0
    b.step(4)
    wb, ub, tb = b.get_state()
    assert tb == ta and bitdiff(wb, wa) == 0 and bitdiff(ub, ua) == 0
    if name == "t2d":
        assert b.stored_rows() == 1  # the restarted axisymmetric state is collapsed again
    a.close()
    b.close()


def test_model_output1_equals_host_writer(tmp_path):
    p = configs.U_PPM.resolved()
    a, g, m = fresh(p)
    a.set_state(native.setup_torus(p, g, m, seed=2, noise=0.02))
    a.step(2)
    a.output1(m["detg"], g["x1"], g["x2"], g["x3"], path=tmp_path / "dev.outdat", nd=5)
    w, _, t = a.get_state()
    native.output1(tmp_path / "host.outdat", w, m["detg"], g["x1"], g["x2"], g["x3"], t, nd=5)
    assert (tmp_path / "dev.outdat").read_bytes() == (tmp_path / "host.outdat").read_bytes()
    a.close()


def run_fortran(p, nsteps, cwd, files):
    cmd = [native.FORTRAN_MAIN, str(p.imax), str(p.jmax), str(p.kmax), str(p.irec), str(p.irkt), repr(p.cfl),
           str(nsteps), "out.bin", "0.02", "0", str(files)]
    print(" ".join(cmd), "in", cwd, flush=True)
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=240)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    raw = open(os.path.join(cwd, "out.bin"), "rb").read()
    ni, nj, nk, ns = (int(x) for x in np.frombuffer(raw[:16], dtype=np.int32))
    off = 16 + 8 * ns
    time = float(np.frombuffer(raw[off:off + 8], dtype=np.float64)[0])
    off += 8
    n = 9 * ni * nj * nk
    uri = np.frombuffer(raw[off:off + 8 * n], dtype=np.float64).reshape((9, ni, nj, nk), order="F")
    uu = np.frombuffer(raw[off + 8 * n:off + 16 * n], dtype=np.float64).reshape((9, ni, nj, nk), order="F")
    return time, uri, uu


def test_fortran_loop_restarts_from_its_own_files(tmp_path):
    if not os.path.exists(native.FORTRAN_MAIN):
        raise native.RaishinError(f"{native.FORTRAN_MAIN} is not built; run raishinsource_amd/_build.py")
    p = configs.U_PPM.resolved()
    straight, split = tmp_path / "straight", tmp_path / "split"
    straight.mkdir()
    split.mkdir()
    t7, w7, u7 = run_fortran(p, 7, straight, 0)
    t3, w3, u3 = run_fortran(p, 3, split, 1)       # 3 steps, then store1 + output1
    r = native.restar2(split / "restart000.outdat", (p.imax, p.jmax, p.kmax))
    assert r["time"] == t3 and bitdiff(r["uri"], w3) == 0 and bitdiff(r["uu"], u3) == 0
    assert (r["it0"], r["ih0"], r["nd"]) == (4, 3, 1)
    g = native.setup_grid(p)
    m = native.setup_metric(p, g)
    native.output1(split / "host.outdat", w3, m["detg"], g["x1"], g["x2"], g["x3"], t3, nd=1)
    assert (split / "structr0000-0001.outdat").read_bytes() == (split / "host.outdat").read_bytes()
    t, w, u = run_fortran(p, 4, split, 3)          # restar2, 4 more steps
    assert t == t7 and bitdiff(w, w7) == 0 and bitdiff(u, u7) == 0


def test_checkpoint_succeeds_with_an_earlier_error_pending(tmp_path):
    """The model entry points report their own outcome: an unrelated sticky
    error left by an earlier call does not make a good checkpoint fail."""
    import ctypes as C
    p = configs.U_PPM.resolved()
    a, g, m = fresh(p)
    a.set_state(native.setup_torus(p, g, m, seed=1, noise=0.02))
    b, _, _ = fresh(p)
    lib = native.lib()
    lib.raishin_clear_error()
    assert lib.raishin_model_tables(a.h, 99, None) == -1  # leaves "bad argument" pending
    assert lib.raishin_model_store1(a.h, str(tmp_path / "r.outdat").encode(), None, None, 1, 0, 0, 0) == 0
    assert lib.raishin_model_restar2(b.h, str(tmp_path / "r.outdat").encode(), None, None, None, None, None, 0) == 0
    buf = C.create_string_buffer(256)
    assert lib.raishin_last_error(buf, 256) == 1 and b"bad argument" in buf.value
    lib.raishin_clear_error()
    wa, ua, ta = a.get_state()
    wb, ub, tb = b.get_state()
    assert ta == tb and bitdiff(wa, wb) == 0 and bitdiff(ua, ub) == 0
    a.close()
    b.close()
