"""The C ABI library without a GPU: it loads, exports every entry point that
include/vitrum_hip.h declares, reports a missing device as a status code,
and turns exceptions inside the library into status codes."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from vitrum_amd import native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "vitrum_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(vt_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_the_proposed_minimum():
    names = set(declared_functions())
    for n in ("vt_create", "vt_set_scene", "vt_render_frame", "vt_capture", "vt_reset", "vt_destroy",
              "vt_trace_closest", "vt_trace_any", "vt_stats", "vt_last_error"):
        assert n in names


def test_library_exports_every_declared_function():
    lib = native.lib()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", native.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    assert set(declared_functions()) <= exported


def test_create_without_a_device_is_a_status_code():
    lib = native.lib()
    h = C.c_void_p()
    st = lib.vt_create(C.c_int32(0), None, C.byref(h))
    if st == 0:  # a GPU is present: nothing to check here
        lib.vt_destroy(h)
        pytest.skip("a HIP device is present")
    assert st == -5 and not h.value  # VT_ERR_NO_DEVICE
    assert "no HIP device" in native.last_error()


def test_exceptions_become_status_codes():
    """An allocation failure inside the builder (forced with the test hook)
    returns VT_ERR_OUT_OF_MEMORY with a message instead of unwinding through
    extern "C"; the next call works normally."""
    tri = {"kind": "mesh", "positions": np.array([0, 0, 0, 1, 0, 0, 0, 1, 0], np.float32),
           "normals": np.tile([0, 0, 1], 3).astype(np.float32), "indices": np.array([0, 1, 2], np.uint32)}
    native.inject_failure(1)
    with pytest.raises(native.VtError) as ei:
        native.pack_scene([tri])
    assert ei.value.status == -7 and "out of host memory" in str(ei.value)
    pack = native.pack_scene([tri])
    assert pack["bvh_nodes"].size == 8
