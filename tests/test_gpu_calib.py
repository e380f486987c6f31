"""The measurement entry points behind bench.py's compute roofline and the
fused kernels' LDS layout (zdt_calib_valu, zdt_calib_lds): they run, return
finite positive rates, and reject arguments outside their ranges."""
import ctypes as C
import math

import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def L(gpu):
    from fdrt_hip import _lib
    return _lib


def test_calib_valu(L):
    r = C.c_double()
    L.check(L.lib().zdt_calib_valu(2, 4, C.byref(r)))
    assert math.isfinite(r.value) and r.value > 1e12  # lane operations per second
    for form, chains in ((4, 1), (0, 3)):
        with pytest.raises(L.ZDTError):
            L.check(L.lib().zdt_calib_valu(form, chains, C.byref(r)))


def test_calib_lds(L):
    r = C.c_double()
    costs = []
    for pat in range(13):
        L.check(L.lib().zdt_calib_lds(pat, C.byref(r)))
        assert math.isfinite(r.value) and r.value > 0.0
        costs.append(r.value)
    print("ns per wavefront access and CU:", [round(c, 3) for c in costs], flush=True)
    for pat in (-1, 13):
        with pytest.raises(L.ZDTError):
            L.check(L.lib().zdt_calib_lds(pat, C.byref(r)))
