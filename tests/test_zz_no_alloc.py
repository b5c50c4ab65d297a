"""raishin_model_cdtcfl4_noalloc: the CFL reduction with its per-workgroup
partials in the library's fixed device array, called through the C-ABI,
against raishin_model_cdtcfl4 (the same kernels on the model's own
buffers), bit for bit: two calls in a row on different models (the second
smaller, with a NaN pressure in one cell, which the reduction's
sequential-MIN semantics pass over), a call after steps, no change in free
device memory across calls, and the size limit's error."""
import ctypes as C
from dataclasses import replace

import numpy as np
import pytest

from raishinsource_amd import configs, native

pytestmark = pytest.mark.gpu


def hip():
    """The HIP runtime the library itself runs on (the test allocates the
    caller's output buffer with it, as a caller would)."""
    native.lib()
    h = C.CDLL("libamdhip64.so.7")
    h.hipMemGetInfo.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    for f in (h.hipMemGetInfo, h.hipMalloc, h.hipFree, h.hipMemcpy):
        f.restype = C.c_int
    return h


class DeviceOut:
    """Four doubles of device memory owned by the caller."""

    def __init__(self, h):
        self.h = h
        self.ptr = C.c_void_p()
        assert h.hipMalloc(C.byref(self.ptr), 32) == 0

    def read(self):
        host = (C.c_double * 4)()
        assert self.h.hipMemcpy(host, self.ptr, 32, 2) == 0  # device to host
        return tuple(host)

    def close(self):
        self.h.hipFree(self.ptr)


def free_bytes(h):
    f, t = C.c_size_t(0), C.c_size_t(0)
    assert h.hipMemGetInfo(C.byref(f), C.byref(t)) == 0
    return f.value


def model(name, nan_at=None):
    p = configs.ALL[name].resolved()
    g = native.setup_grid(p)
    m = native.setup_metric(p, g)
    uri = native.setup_torus(p, g, m, seed=0, noise=0.02)
    if nan_at is not None:
        uri[(4,) + nan_at] = np.nan
    mod = native.Model(p, g, m)
    mod.set_state(uri)
    return mod


def last_error():
    buf = C.create_string_buffer(512)
    native.lib().raishin_last_error(buf, 512)
    return buf.value.decode()


def noalloc(mod, out):
    rc = native.lib().raishin_model_cdtcfl4_noalloc(C.c_void_p(mod.h), out.ptr)
    assert rc == 0, last_error()
    mod.sync()
    return out.read()


def same_bits(a, b):
    return np.array(a, dtype=np.float64).view(np.uint64).tolist() == np.array(b, dtype=np.float64).view(
        np.uint64).tolist()


def test_cdtcfl4_noalloc_matches_cdtcfl4():
    h = hip()
    out = DeviceOut(h)
    big = model("t3dppm")
    small = model("uppm", nan_at=(10, 5, 9))
    try:
        r1 = noalloc(big, out)  # the first call may load the reduction's code
        e1 = big.cdtcfl4()
        print("t3dppm", r1, e1)
        assert same_bits(r1, e1) and np.isfinite(r1).all()
        f0 = free_bytes(h)
        r2 = noalloc(small, out)
        assert free_bytes(h) >= f0  # no device allocation in the call
        e2 = small.cdtcfl4()
        print("uppm with a NaN", r2, e2)
        assert same_bits(r2, e2) and not same_bits(r2, r1)
        big.step(2)
        e3 = big.cdtcfl4()
        f2 = free_bytes(h)
        r3 = noalloc(big, out)
        assert free_bytes(h) >= f2
        assert same_bits(r3, e3)
    finally:
        big.close()
        small.close()
        out.close()


def test_cdtcfl4_noalloc_rejects_models_over_its_limit():
    out = DeviceOut(hip())
    p = replace(configs.C2_PPM, name="over", imax=530, jmax=256, kmax=256, ymax=None).resolved()
    assert ((p.imax - 2) * (p.jmax - 2) * (p.kmax - 2) + 255) // 256 > 131072
    g = native.setup_grid(p)
    mod = native.Model.from_grid(p, g)
    try:
        rc = native.lib().raishin_model_cdtcfl4_noalloc(C.c_void_p(mod.h), out.ptr)
        assert rc == -1
        msg = last_error()
        assert "over the fixed scratch" in msg, msg
        native.lib().raishin_clear_error()
    finally:
        mod.close()
        out.close()
