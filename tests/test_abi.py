"""The C-ABI library loads and exports every function the public headers
declare, and its host-only entry points reject bad arguments (CPU only:
nothing here touches a GPU)."""
import ctypes
import os
import re

import pytest

import rtamd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = ["rt_api.h", "rt_scenes.h", "rt_io.h", "rt_debug.h"]


def declared_functions():
    names = set()
    for h in HEADERS:
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(rt_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_headers_declare_the_expected_surface():
    names = declared_functions()
    for must in ("rt_bvh_build", "rt_bvh_refit", "rt_trace", "rt_render", "rt_batch_render", "rt_diff_forward",
                 "rt_diff_backward", "rt_obj_load", "rt_write_bmp", "rt_resolve_supersample"):
        assert must in names
    assert len(names) >= 35


def test_library_exports_every_declared_function():
    L = rtamd.lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing
    # the Python binding's list is a subset of the headers
    assert set(rtamd.EXPORTED) <= set(declared_functions())


def test_host_only_entry_points_reject_bad_arguments(tmp_path):
    L = rtamd.lib()
    L.rt_version.restype = ctypes.c_char_p
    assert L.rt_version()
    L.rt_last_error.restype = ctypes.c_char_p
    assert L.rt_last_error(None) == b"null context"
    h = ctypes.c_void_p()
    err = ctypes.create_string_buffer(128)
    rtamd._bind_io(L)
    assert L.rt_obj_load(None, None, ctypes.byref(h), err, 128) == rtamd.RT_ERR_INVALID
    assert L.rt_obj_load(str(tmp_path / "none.obj").encode(), None, ctypes.byref(h), err, 128) == rtamd.RT_ERR_INVALID
    assert b"cannot open" in err.value and not h.value
    assert L.rt_write_bmp(str(tmp_path / "x.bmp").encode(), 0, 4, None) == rtamd.RT_ERR_INVALID
    from rtamd import diff
    diff._bind_diff(L)
    cam = rtamd.camera(0, 16)
    assert L.rt_diff_record_bytes(ctypes.byref(cam), ctypes.byref(rtamd.render_opts(1))) == 0
    cam = rtamd.camera(16, 8)
    assert L.rt_diff_record_bytes(ctypes.byref(cam), ctypes.byref(rtamd.render_opts(3, depth=2))) == 16 * 8 * 3 * 3 * 8
