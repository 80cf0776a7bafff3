"""The kernels convert f32 to integers with one GPU instruction (v_cvt_i32_f32 /
v_cvt_u32_f32) plus an integer clamp; the oracle uses the portable definitions
in include/st_math.h.  This runs both forms on the device over all 2^32 bit
patterns of the input and requires them to agree everywhere (NaN, infinities,
the saturation ends, denormals included)."""
import ctypes as C
import os

import pytest

pytestmark = pytest.mark.gpu

LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c", "libconvert_check.so")


def test_device_conversions_match_portable():
    if not os.path.isfile(LIB):
        raise RuntimeError(f"{LIB} not built (run __graft_entry__.build())")
    lib = C.CDLL(LIB)
    bad, first = C.c_ulonglong(0), C.c_uint32(0)
    assert lib.st_check_conversions(C.byref(bad), C.byref(first)) == 0
    assert bad.value == 0, f"{bad.value} mismatches, first at bits {first.value:#010x}"
