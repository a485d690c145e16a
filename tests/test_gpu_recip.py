"""recip_rn (venusaur_amd/csrc/vz_device.h) equals the IEEE reciprocal.

The step kernels compute a ray's 1/d, the hit normal's 1/r and normalise
directions with recip_rn: v_rcp_f32 refined by one fma Newton-Raphson step
for |x| in [2^-100, 2^100), the IEEE division elsewhere.  The oracle divides,
so images and node-visit counts stay bit-exact only if the two agree on every
input of that range.  This runs the product's recip_rn over all of them
(3,355,443,200 floats) and also checks that the comparison can fail: the
unrefined v_rcp_f32 must differ somewhere.
"""
import ctypes as C
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "gpu_checks", "librecipcheck.so")


def _lib():
    if not os.path.exists(LIB):
        subprocess.run(["make", "-C", os.path.join(HERE, "gpu_checks")], check=True)
    lib = C.CDLL(LIB)
    lib.vz_check_recip_rn.restype = C.c_int
    lib.vz_check_recip_rn.argtypes = [C.POINTER(C.c_ulonglong), C.POINTER(C.c_uint32)]
    return lib


@pytest.mark.gpu
def test_recip_rn_equals_ieee_reciprocal_on_its_whole_range():
    print("case: recip_rn against 1.0f / x, |x| in [2^-100, 2^100), both signs", flush=True)
    out = (C.c_ulonglong * 5)()
    bad = (C.c_uint32 * 4)()
    assert _lib().vz_check_recip_rn(out, bad) == 0
    ref_diff, rn_diff, raw_diff, seen, expected = list(out)
    assert expected == 2 * 200 * (1 << 23) and seen == expected
    assert ref_diff == 0, "the division and the double-rounded reference disagree"
    assert rn_diff == 0, f"recip_rn differs from 1.0f / x for {rn_diff} inputs, e.g. {[hex(b) for b in bad]}"
    assert raw_diff > 0, "the unrefined v_rcp_f32 never differed: the comparison checks nothing"
