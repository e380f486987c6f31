"""C-ABI checks that need no GPU: the library loads, exports every symbol
include/zdt.h declares, and its host-only entry points (sizing, argument
validation) agree with the oracle."""
import ctypes as C
import re
from pathlib import Path

import pytest

from oracle import host_oracle as H

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "zdt.h"


def declared_symbols():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(zdt_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def L():
    from fdrt_hip import _lib
    return _lib.lib()


def test_header_declares_entry_points():
    syms = declared_symbols()
    for s in ["zdt_create", "zdt_run", "zdt_input", "zdt_output", "zdt_findprotohits", "zdt_fdrstats"]:
        assert s in syms


def test_library_exports_every_declared_symbol(L):
    for s in declared_symbols():
        assert hasattr(L, s), f"{s} missing from libzdt_hip.so"


def test_bindings_cover_every_symbol():
    from fdrt_hip import _lib
    assert sorted(_lib.SIGNATURES) == declared_symbols()


def _f(vals):
    return (C.c_int * len(vals))(*vals), len(vals)


@pytest.mark.parametrize("Nt,Nr,factors", [(2, 3, (2, 3, 5)), (16, 129, (2, 3, 5)), (16, 129, (2,)),
                                           (60, 481, (2, 3, 5)), (64, 1025, (2, 3, 5)), (7, 11, (2, 3, 5, 7))])
def test_calc_nl_grow_nr(L, Nt, Nr, factors):
    arr, n = _f(factors)
    assert L.zdt_calc_nl(Nt, Nr, arr, n) == H.calcNl(Nt, Nr, factors)
    assert L.zdt_grow_nr(Nt, Nr, arr, n) == H.growNr(Nt, Nr, factors)
    assert L.zdt_calc_nl(Nt, Nr, None, 0) == H.calcNl(Nt, Nr)


@pytest.mark.parametrize("args", [(1 << 20, 16, 129, 1, 1), (65536, 16, 33, 2, 3), (3, 2, 3, 1, 1)])
def test_estimate_memory(L, args):
    assert L.zdt_estimate_memory(*args, None, 0) == H.estimate_memory(*args)


def test_python_sizing_wrappers():
    import fdrt_hip as F
    assert F.calcNl(16, 129) == 144
    assert F.growNr(16, 100) == H.growNr(16, 100)
    assert F.estimate_memory(1 << 20, 16, 129) == H.estimate_memory(1 << 20, 16, 129)


def test_bad_factor_rejected(L):
    arr, n = _f((1, 2))
    assert L.zdt_calc_nl(16, 129, arr, n) == -1


def test_create_validation_errors_before_device(L):
    from fdrt_hip import _lib
    h = C.c_void_p()
    rc = L.zdt_create(None, 0, 16, 0.0, 1.0, 10, None, 0, 0, 0, None, C.byref(h))
    assert rc == _lib.ZDT_EINVAL
    assert "invalid shape" in L.zdt_last_error().decode()
    rc = L.zdt_create(None, 17 * 4, 16, 0.0, 1.0, 10, None, 0, 0, 0, None, C.byref(h))
    assert rc == _lib.ZDT_ENOTSUP
    assert "prime factor" in L.zdt_last_error().decode()
    arr, n = _f((0,))
    rc = L.zdt_create(None, 64, 16, 0.0, 1.0, 10, arr, n, 0, 0, None, C.byref(h))
    assert rc == _lib.ZDT_EINVAL
    rc = L.zdt_create(None, 64, 16, float("nan"), 1.0, 10, None, 0, 0, 0, None, C.byref(h))
    assert rc == _lib.ZDT_EINVAL
    rc = L.zdt_create(None, 64, 16, 0.0, 1.0, 10, None, 0, 0, 0, None, None)
    assert rc == _lib.ZDT_EINVAL
    rc = L.zdt_create(None, 1 << 16, 16, 0.0, 1.0, 10, None, 0, 0, 7, None, C.byref(h))
    assert rc == _lib.ZDT_EINVAL
    assert "precision" in L.zdt_last_error().decode()


def test_null_workspace_errors(L):
    from fdrt_hip import _lib
    assert L.zdt_run(None, 0.0, 0, None, None) == _lib.ZDT_EINVAL
    assert L.zdt_destroy(None) == _lib.ZDT_OK
    assert L.zdt_fdrstats(None, 0, 10, 10, 10, None, None, None) == _lib.ZDT_EINVAL
    assert L.zdt_path(None, None, None) == _lib.ZDT_EINVAL
    assert L.zdt_version() == 2


def test_plan_validation_errors_before_device(L):
    from fdrt_hip import _lib
    h = C.c_void_p()
    one = (C.c_int64 * 1)(1 << 16)
    f = (C.c_float * 1)(0.0)
    assert L.zdt_plan_ragged(0, one, one, f, f, one, 0, None, C.byref(h)) == _lib.ZDT_EINVAL
    assert L.zdt_plan_ragged(1, one, one, f, f, one, 0x4, None, C.byref(h)) == _lib.ZDT_EINVAL
    assert "flags" in L.zdt_last_error().decode()
    assert L.zdt_plan_ragged(1, None, one, f, f, one, 0, None, C.byref(h)) == _lib.ZDT_EINVAL
    assert L.zdt_plan_launch(None) == _lib.ZDT_EINVAL
    assert L.zdt_plan_info(None, 0, None, None, None, None) == _lib.ZDT_EINVAL
    assert L.zdt_plan_bind(None, None, None, None, 0, None, 0, None, None) == _lib.ZDT_EINVAL
    assert L.zdt_plan_destroy(None) == _lib.ZDT_OK


def test_batchrates_mirror_matches_oracle():
    import fdrt_hip as F
    for args, kw in [((16, 0.1, 4.0), {"Nrb": 100}), ((16, 0.1, 8.0), {"Nrb": 50}),
                     ((16, -0.1, 8.0), {"Nrb": 50}), ((4, 1.0, 5.0), {"Nrb": 100}),
                     ((60, 0.05, -3.0, 2.0), {"Nrb": 200}), ((16, 0.25, 6.0), {})]:
        a = F.batchrates(*args, **kw)
        b = H.batchrates(*args, **kw)
        assert [(x.first, x.step, x.length) for x in a] == b


def test_rates_params():
    from fractions import Fraction

    import numpy as np

    from fdrt_hip.zdtutils import RateRange, rates_params
    assert rates_params(range(-1, 2)) == (-1, 1, 3)
    r0, dr, nr = rates_params(np.arange(-4, 4.0001, 0.0625))
    assert (r0, dr, nr) == (np.float32(-4), np.float32(0.0625), 129)
    r0, dr, nr = rates_params(RateRange(Fraction(-8, 3), Fraction(1, 15), 81))
    assert r0 == np.float32(np.float32(-8) / np.float32(3)) and nr == 81
    with pytest.raises(ValueError, match="evenly spaced"):
        rates_params([0.0, 1.0, 3.0])
