"""The allocation-free findprotohits (zdt_findprotohits_noalloc) through the
C-ABI against zdt_findprotohits (src/fdrutils.jl:148-169): identical hit
lists, counts and truncation on two calls in a row, the second smaller than
the first (the shared device array keeps the first call's contents), and the
size limit rejected before any launch."""
import ctypes as C
import sys

import numpy as np
import pytest

from oracle import synth

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def F(gpu):
    import fdrt_hip
    return fdrt_hip


def hits_via(L, fn, fdrs, Nf, Nr, thr, cap):
    import torch
    out = torch.full((max(cap, 1),), -1, dtype=torch.int64, device=fdrs[0].device)
    arr = (C.c_void_p * len(fdrs))(*[d.data_ptr() for d in fdrs])
    n = C.c_int64(-1)
    rc = getattr(L, fn)(arr, len(fdrs), Nf, Nr, Nf, thr, out.data_ptr() if cap > 0 else None, cap, C.byref(n), None)
    torch.cuda.synchronize()
    return rc, n.value, out.cpu().numpy()[:min(cap, max(n.value, 0))]


def test_noalloc_matches_findprotohits_two_calls(F):
    from fdrt_hip import _lib
    L = _lib.lib()
    cases = [
        # (Nf, Nt, Nr, batches): three 2^18 x 33 FDRs, then one 2^16 x 9
        (1 << 18, 16, 33, 3),
        (1 << 16, 8, 9, 1),
    ]
    for Nf, Nt, Nr, nb in cases:
        s = synth.spectrogram(Nf, Nt, [(Nf * 0.3, 1.25, 30.0), (Nf * 0.6, -0.5, 25.0)], seed=Nf % 991 + nb)
        ws = F.ZDTWorkspace(s, -2.0, 0.125, Nr)
        fdrs = [F.zdtfdr(ws, r0=-2.0 + 0.125 * Nr * b) for b in range(nb)]
        thr = F.fdrdenormalize(6.0, fdrs)
        say(f"noalloc Nf={Nf} Nr={Nr} nb={nb} thr={thr:.4f}")
        rc0, n0, h0 = hits_via(L, "zdt_findprotohits", fdrs, Nf, Nr, thr, 1 << 16)
        rc1, n1, h1 = hits_via(L, "zdt_findprotohits_noalloc", fdrs, Nf, Nr, thr, 1 << 16)
        assert rc0 == rc1 == 0 and n0 == n1 > 0
        np.testing.assert_array_equal(h0, h1)
        # truncation and counting only, as zdt_findprotohits does them
        rc2, n2, h2 = hits_via(L, "zdt_findprotohits_noalloc", fdrs, Nf, Nr, thr, 5)
        assert rc2 == _lib.ZDT_ETRUNC and n2 == n0
        np.testing.assert_array_equal(h2, h0[:5])
        rc3, n3, _ = hits_via(L, "zdt_findprotohits_noalloc", fdrs, Nf, Nr, thr, 0)
        assert rc3 == _lib.ZDT_ETRUNC and n3 == n0
        ws.close()


def test_noalloc_rejects_sizes_beyond_its_array(F):
    import torch
    from fdrt_hip import _lib
    L = _lib.lib()
    d = torch.zeros(16, device="cuda")
    arr = (C.c_void_p * 1)(d.data_ptr())
    n = C.c_int64(-1)
    # 2^30 x 64 cells = 2^24 tiles, past the 2^21-tile array: refused on the host
    rc = L.zdt_findprotohits_noalloc(arr, 1, 1 << 30, 64, 1 << 30, 1.0, None, 0, C.byref(n), None)
    assert rc == _lib.ZDT_ENOTSUP and n.value == -1
    assert b"limit" in L.zdt_last_error()
    # 8 matrices of 2^20 x 1025 (2^18 + 256 tiles each) also exceed it together
    arr8 = (C.c_void_p * 8)(*([d.data_ptr()] * 8))
    rc = L.zdt_findprotohits_noalloc(arr8, 8, 1 << 20, 1025, 1 << 20, 1.0, None, 0, C.byref(n), None)
    assert rc == _lib.ZDT_ENOTSUP
