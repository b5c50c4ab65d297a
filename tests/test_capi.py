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


@pytest.mark.parametrize("header", ["raishin_hip.h", "raishin_setup.h", "raishin_ref.h", "raishin_io.h"])
def test_library_exports_header(header):
    lib = C.CDLL(native.LIB_PATH)
    names = declared(header)
    assert names, header
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_exports():
    names = set(declared("raishin_hip.h")) | set(declared("raishin_setup.h")) | set(declared("raishin_io.h"))
    assert names <= set(native.EXPORTS), names - set(native.EXPORTS)
    assert set(declared("raishin_ref.h")) == set(native.REF_EXPORTS)


def test_fortran_interfaces_bind_every_reference_entry_point():
    """raishinsource_amd/fortran/raishin_c.f90 declares a bind(C) interface for
    every reference-named entry point, and the wrappers forward each one."""
    import re as _re
    fdir = os.path.join(ROOT, "raishinsource_amd", "fortran")
    iface = open(os.path.join(fdir, "raishin_c.f90")).read()
    bound = set(_re.findall(r'bind\(C, name="(raishin_\w+)"\)', iface))
    assert set(native.REF_EXPORTS) <= bound, set(native.REF_EXPORTS) - bound
    wrappers = open(os.path.join(fdir, "raishin_wrappers.f90")).read()
    called = set(_re.findall(r"call (raishin_\w+)\(", wrappers))
    routines = set(native.REF_EXPORTS) - {"raishin_init", "raishin_geometry", "raishin_finalize", "raishin_push",
                                          "raishin_pull", "raishin_sync", "raishin_get_counters",
                                          "raishin_reset_counters"}
    io = {"raishin_store1", "raishin_restar2", "raishin_output1"}
    assert io <= bound and io <= called, io - (bound & called)
    called -= io | {"raishin_pull", "raishin_push"}
    assert routines == called, routines ^ called


def test_reference_entry_points_report_order_errors_without_gpu_work():
    """A reference-named call before raishin_geometry, and a configuration the
    library does not build, report through the sticky error."""
    lib = native.lib()
    buf = C.create_string_buffer(256)
    lib.raishin_clear_error()
    one = C.c_int(1)
    lib.raishin_ident(None, None, C.byref(one), C.byref(one), C.byref(one), C.byref(one), C.byref(one),
                      C.byref(one))
    assert lib.raishin_last_error(buf, 256) == 1 and b"raishin_geometry" in buf.value
    lib.raishin_clear_error()
    from dataclasses import replace
    from raishinsource_amd import configs
    ps = native.pram_struct(replace(configs.G_MP5, ieos=3))
    lib.raishin_init(C.byref(ps))
    assert lib.raishin_last_error(buf, 256) == 1 and b"ieos=3" in buf.value
    lib.raishin_clear_error()


def test_error_path_without_gpu_work():
    """A rejected configuration reports through the sticky error, not a crash."""
    from dataclasses import replace
    from raishinsource_amd import configs
    p = replace(configs.G_MP5, irec=2)   # minmod limiter: not built
    ps = native.pram_struct(p)
    lib = native.lib()
    buf = C.create_string_buffer(256)
    lib.raishin_clear_error()
    z = (C.c_double * 4)()
    h = lib.raishin_model_create(C.byref(ps), *([C.cast(z, C.c_void_p)] * 19))
    assert not h
    assert lib.raishin_last_error(buf, 256) == 1
    assert b"irec=2" in buf.value
    lib.raishin_clear_error()
