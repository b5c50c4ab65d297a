"""The restart and analysis files (restar.f90 store1 / restar2 / output1)
against flang's own unformatted sequential I/O: the library's C writer
produces the bytes a Fortran WRITE of the same records produces, and each
side reads what the other wrote.  Host code only; the device-resident
checkpoint round trip is in test_gpu_files.py."""
import os
import subprocess

import numpy as np
import pytest

from raishinsource_amd import native

HERE = os.path.dirname(os.path.abspath(__file__))
FC = os.environ.get("FC_FLANG", "/opt/rocm/lib/llvm/bin/flang")
SHAPE = (7, 5, 4)


@pytest.fixture(scope="module")
def frec(tmp_path_factory):
    if not os.path.exists(FC):
        pytest.skip("flang absent")
    exe = str(tmp_path_factory.mktemp("frec") / "fortran_records")
    subprocess.run([FC, "-O0", os.path.join(HERE, "golden", "fortran_records.f90"), "-o", exe], check=True,
                   capture_output=True)
    return exe


def state(seed):
    rng = np.random.default_rng(seed)
    a = {k: np.asfortranarray(rng.standard_normal((9,) + SHAPE)) for k in ("uu", "u0", "uri", "uri0")}
    a["uri"][0, 1, 2, 3] = np.nan
    a["uu"][4, 0, 0, 0] = -0.0
    a.update(myrank=3, nd=12, it0=4567, ih0=89, time=float(rng.random() * 100))
    return a


def state_bin(path, a):
    with open(path, "wb") as f:
        f.write(np.array([a["myrank"], a["nd"], a["it0"], a["ih0"]], dtype="<i4").tobytes())
        f.write(np.array([a["time"]], dtype="<f8").tobytes())
        for k in ("uu", "u0", "uri", "uri0"):
            f.write(a[k].tobytes(order="F"))


def bits(x):
    return np.ascontiguousarray(x).view(np.uint64)


def test_store1_bytes_equal_fortran_write(frec, tmp_path):
    a = state(0)
    c_file, f_file = tmp_path / "c.outdat", tmp_path / "f.outdat"
    native.store1(c_file, a["uu"], a["u0"], a["uri"], a["uri0"], a["it0"], a["ih0"], a["time"], a["nd"],
                  myrank=a["myrank"])
    state_bin(tmp_path / "s.bin", a)
    subprocess.run([frec, "w", *map(str, SHAPE), str(tmp_path / "s.bin"), str(f_file)], check=True)
    cb, fb = c_file.read_bytes(), f_file.read_bytes()
    assert len(cb) == 48 + 320 * np.prod(SHAPE)
    assert cb == fb


def test_restart_files_read_both_ways(frec, tmp_path):
    a = state(1)
    state_bin(tmp_path / "s.bin", a)
    f_file = tmp_path / "f.outdat"
    subprocess.run([frec, "w", *map(str, SHAPE), str(tmp_path / "s.bin"), str(f_file)], check=True)
    r = native.restar2(f_file, SHAPE, myrank=a["myrank"])
    for k in ("uu", "u0", "uri", "uri0"):
        assert np.array_equal(bits(r[k]), bits(a[k])), k
    assert (r["it0"], r["ih0"], r["nd"], r["time"]) == (a["it0"], a["ih0"], a["nd"], a["time"])
    c_file = tmp_path / "c.outdat"
    native.store1(c_file, a["uu"], a["u0"], a["uri"], a["uri0"], a["it0"], a["ih0"], a["time"], a["nd"],
                  myrank=a["myrank"])
    subprocess.run([frec, "r", *map(str, SHAPE), str(c_file), str(tmp_path / "back.bin")], check=True)
    assert (tmp_path / "back.bin").read_bytes() == (tmp_path / "s.bin").read_bytes()


def test_output1_bytes_equal_fortran_write(frec, tmp_path):
    rng = np.random.default_rng(2)
    ni, nj, nk = SHAPE
    x = [rng.random(n) for n in SHAPE]
    uri = np.asfortranarray(rng.standard_normal((9,) + SHAPE))
    detg = np.asfortranarray(rng.random(SHAPE))
    t = 7.25
    native.output1(tmp_path / "c.outdat", uri, detg, *x, t, nd=3, myrank=0, npe=1)
    with open(tmp_path / "d.bin", "wb") as f:
        f.write(np.array([t], "<f8").tobytes() + np.array([0, 1], "<i4").tobytes())
        for a in (*x, uri, detg):
            f.write(np.asfortranarray(a).tobytes(order="F"))
    subprocess.run([frec, "o", *map(str, SHAPE), str(tmp_path / "d.bin"), str(tmp_path / "f.outdat")], check=True)
    cb = (tmp_path / "c.outdat").read_bytes()
    assert len(cb) == 24 + 32 + 112 * ni * nj * nk
    assert cb == (tmp_path / "f.outdat").read_bytes()
    # the i-outer, k-inner order (restar.f90:171-179): the second cell record is (i=1, j=1, k=2)
    rec = np.frombuffer(cb[56 + 112 + 4:56 + 112 + 4 + 104], "<f8")
    assert rec[0] == x[0][0] and rec[1] == x[1][0] and rec[2] == x[2][1]
    assert np.array_equal(rec[3:12], uri[:, 0, 0, 1]) and rec[12] == detg[0, 0, 1]


def test_reference_file_names(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    a = state(3)
    native.store1(None, a["uu"], a["u0"], a["uri"], a["uri0"], 1, 2, 0.5, 0, myrank=7)
    assert (tmp_path / "restart007.outdat").exists()
    native.output1(None, a["uri"], np.asfortranarray(np.ones(SHAPE)), *[np.zeros(n) for n in SHAPE], 0.5, nd=12,
                   myrank=0)
    assert (tmp_path / "structr0000-0012.outdat").exists()


def test_bad_restart_files_are_rejected(tmp_path):
    a = state(4)
    good = tmp_path / "g.outdat"
    native.store1(good, a["uu"], a["u0"], a["uri"], a["uri0"], 1, 2, 0.5, 0, myrank=0)
    data = bytearray(good.read_bytes())
    with pytest.raises(native.RaishinError, match="written by rank 0, read by rank 1"):
        native.restar2(good, SHAPE, myrank=1)
    with pytest.raises(native.RaishinError, match="bytes"):
        native.restar2(good, (7, 5, 5))
    (tmp_path / "short.outdat").write_bytes(bytes(data[:-8]))
    with pytest.raises(native.RaishinError, match="bytes"):
        native.restar2(tmp_path / "short.outdat", SHAPE)
    bad = bytearray(data)
    bad[48 + 320 * 5 + 80] ^= 1  # length word of cell 5's second record
    (tmp_path / "bad.outdat").write_bytes(bytes(bad))
    with pytest.raises(native.RaishinError, match="cell 5"):
        native.restar2(tmp_path / "bad.outdat", SHAPE)
    with pytest.raises(native.RaishinError, match="open"):
        native.restar2(tmp_path / "missing.outdat", SHAPE)


def test_null_arguments_are_errors():
    import ctypes as C
    lib = native.lib()
    lib.raishin_clear_error()
    one = C.byref(C.c_int(1))
    lib.raishin_store1(None, None, None, None, None, one, one, C.byref(C.c_double(0)), one, one, one, one, one, one,
                       one, one, one)
    buf = C.create_string_buffer(256)
    assert lib.raishin_last_error(buf, 256) == 1 and b"null argument" in buf.value
    lib.raishin_clear_error()
    assert lib.raishin_model_store1(None, None, None, None, 0, 0, 0, 0) == -1
    lib.raishin_clear_error()
