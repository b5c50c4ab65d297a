"""The C-ABI library loads and exports every function include/*.h declares
(no GPU work is started)."""
import ctypes as C
import os
import re

import pytest

from raishinsource_amd import native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w \*]*?\b(raishin_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


@pytest.mark.parametrize("header", ["raishin_hip.h", "raishin_setup.h"])
def test_library_exports_header(header):
    lib = C.CDLL(native.LIB_PATH)
    names = declared(header)
    assert names, header
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_exports():
    names = set(declared("raishin_hip.h")) | set(declared("raishin_setup.h"))
    assert names <= set(native.EXPORTS) | {"raishin_init", "raishin_cdtcfl4"}, names - set(native.EXPORTS)


def test_error_path_without_gpu_work():
    """A rejected configuration reports through the sticky error, not a crash."""
    from dataclasses import replace
    from raishinsource_amd import configs
    p = replace(configs.G_MP5, irec=1)   # MC limiter: not built
    ps = native.pram_struct(p)
    lib = native.lib()
    buf = C.create_string_buffer(256)
    lib.raishin_clear_error()
    z = (C.c_double * 4)()
    h = lib.raishin_model_create(C.byref(ps), *([C.cast(z, C.c_void_p)] * 19))
    assert not h
    assert lib.raishin_last_error(buf, 256) == 1
    assert b"irec=1" in buf.value
    lib.raishin_clear_error()
