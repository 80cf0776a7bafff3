"""CPU checks of the C ABI library: it loads, and it exports every function
include/splashtank.h declares (no compute call; there is no GPU here)."""
import ctypes
import os
import re

import pytest

from splashtank_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    src = open(os.path.join(ROOT, "include", "splashtank.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(st_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_the_boundary():
    names = header_functions()
    for must in ["st_create", "st_destroy", "st_step", "st_stage", "st_upload_particles", "st_read_particles",
                 "st_set_colliders", "st_read_ctl", "st_dump", "st_grow"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    if not os.path.isfile(capi.LIB_PATH):
        pytest.fail("libsplashtank_hip.so is not built; run __graft_entry__.build()")
    lib = ctypes.CDLL(capi.LIB_PATH)
    missing = [n for n in header_functions() if not hasattr(lib, n)]
    assert not missing, missing
    assert set(header_functions()) <= set(capi.PROTOS)


def test_params_struct_layout_matches_header():
    # sizes the header implies: 7 floats + int + 3 + 4 floats + 4 u32 + u32 + f + 4 u32 + 64 f + u32
    # + 16 whitewater words + 3 drip words + 5 ocean words and 32 waves of 8 floats
    assert ctypes.sizeof(capi.StParams) == 4 * (7 + 1 + 3 + 4 + 4 + 1 + 1 + 1 + 3 + 64 + 1 + 16 + 3 + 5 + 256 + 2)
    assert ctypes.sizeof(capi.StOpts) == 8 * 5 + 4 * 4
    assert ctypes.sizeof(capi.StDims) == 4 * (10 + 6 + 5)


def test_defaults_match_reference():
    lib = capi.load()
    p = capi.StParams()
    lib.st_params_default(ctypes.byref(p))
    q = capi.make_params()
    for f in ("flip", "comp", "vmax", "foam_tau", "foam_sens", "g"):
        assert abs(getattr(p, f) - getattr(q, f)) < 1e-7, f
    assert p.omega == 0.0 and p.tear == 0.0


def test_create_without_gpu_fails_loudly():
    lib = capi.load()
    h = ctypes.c_void_p()
    o = capi.make_opts(1 / 32, (1, 1, 1), max_n=10)
    rc = lib.st_create(ctypes.byref(o), ctypes.byref(h))
    if rc == 0:
        lib.st_destroy(h)
        pytest.skip("a HIP device is present")
    assert rc < 0 and lib.st_last_error()
