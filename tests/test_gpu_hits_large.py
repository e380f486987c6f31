"""GPU fdrstats / findprotohits on matrices large enough for the multi-chunk
tile-count scan (more than 8192 tiles of 4096 cells), on the contiguous
16-byte-aligned path (float4 loads), on a contiguous matrix whose columns are
not 16-byte aligned, on a column-padded matrix (scalar loads) and across
several matrices, against the oracle (src/fdrutils.jl:39-50, :148-169)."""
import sys

import numpy as np
import pytest

from oracle import host_oracle as H

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def F(gpu):
    import fdrt_hip
    return fdrt_hip


def fdr_like(Nf, Nr, pad, seed):
    """(host view, device view) of an Nf x Nr column-major matrix of
    Exponential(1) values with column stride Nf + pad."""
    import torch
    rng = np.random.default_rng(seed)
    a = rng.exponential(1.0, size=(Nr, Nf + pad)).astype(np.float32)
    d = torch.from_numpy(a).cuda()
    return a[:, :Nf].T, d.t()[:Nf, :]


@pytest.mark.parametrize("Nf,Nr,pad", [(1 << 20, 129, 0), ((1 << 20) + 3, 40, 0), (300001, 120, 5)])
def test_stats_hits_large(F, Nf, Nr, pad):
    say(f"stats/hits Nf={Nf} Nr={Nr} pad={pad}")
    h, d = fdr_like(Nf, Nr, pad, seed=Nf % 1009 + Nr)
    assert d.stride(1) == Nf + pad
    m, s = F.fdrstats(d)
    mo, so = H.fdrstats(h)
    assert abs(m - mo) <= 1e-12 * abs(mo) and abs(s - so) <= 1e-12 * so
    hits = F.findprotohits(d, 5.0, snr=True, capacity=1 << 22)
    ref = H.findprotohits(h, 5.0, snr=True)
    assert len(ref) > 50000
    np.testing.assert_array_equal(hits, ref)


def test_hits_large_multimatrix(F):
    say("stats/hits over three 2^20 x 40 matrices")
    pairs = [fdr_like(1 << 20, 40, 0, seed=70 + b) for b in range(3)]
    hs, ds = [p[0] for p in pairs], [p[1] for p in pairs]
    m, s = F.fdrstats(ds)
    mo, so = H.fdrstats(hs)
    assert abs(m - mo) <= 1e-12 * abs(mo) and abs(s - so) <= 1e-12 * so
    hits = F.findprotohits(ds, 4.5, snr=True, capacity=1 << 22)
    ref = H.findprotohits(hs, 4.5, snr=True)
    np.testing.assert_array_equal(hits, ref)
    # an absolute threshold with no hits at all
    assert F.findprotohits(ds, 1e9).shape == (0, 2)
