"""GPU proto-hit clustering (zdt_clusterprotohits; fdrt_hip.clusterprotohits
and findhits) against the oracle (oracle/cluster_oracle.py), exactly: every
hit's peak cell, peak value, size, bounding box, and every proto-hit's hit
number.  The reference returns proto-hits and leaves their clustering to the
next stage (src/fdrutils.jl:141-144, README.md:48-53).

Small matrices cover the radii, several hcat'd batches with hits across the
batch boundary, one hit holding every cell, the column boundary, empty and
single inputs and the C-ABI's errors.  At the full 2^20 x 129 size the SNR-8
hits of the 16-tone spectrogram are checked against the oracle, and denser
thresholds (up to half of the 135M cells, one component spanning the matrix)
against scipy.ndimage (label, find_objects, maximum, sum) on the dense mask."""
import ctypes as C
import json
import sys
from pathlib import Path

import numpy as np
import pytest

from oracle import cluster_oracle as CO
from oracle import host_oracle as H
from oracle import synth

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def F(gpu):
    import fdrt_hip
    return fdrt_hip


def matrices(Nf, Nr, nb, seed, blobs=6):
    """nb (host, device) Nf x Nr column-major Exponential(1) matrices with a
    few bright diagonal streaks, one of them across each batch boundary."""
    import torch
    rng = np.random.default_rng(seed)
    full = rng.exponential(1.0, size=(Nf, Nr * nb)).astype(np.float32)
    for q in range(blobs + nb - 1):
        J0 = (q + 1) * Nr - 2 if q < nb - 1 else int(rng.integers(0, Nr * nb - 4))
        i0 = int(rng.integers(5, Nf - 40))
        slope = int(rng.integers(-3, 4))
        for t in range(4):
            full[i0 + slope * t: i0 + slope * t + 3, J0 + t] += 8.0
    host = [full[:, b * Nr:(b + 1) * Nr].copy() for b in range(nb)]
    dev = [torch.from_numpy(np.ascontiguousarray(h.T)).cuda().t() for h in host]
    return host, dev


def same(hits, ref):
    assert np.array_equal(hits.size, ref["size"])
    assert np.array_equal(hits.peak, ref["peak"])
    assert np.array_equal(hits.value, ref["value"])
    assert np.array_equal(hits.box, ref["box"])
    assert np.array_equal(hits.labels, ref["labels"])


@pytest.mark.parametrize("nb", [1, 3])
@pytest.mark.parametrize("di,dj", [(1, 1), (0, 0), (2, 1), (1, 3), (3, 0)])
@pytest.mark.parametrize("thr", [6.0, 2.5, 0.7])
def test_small_vs_oracle(F, nb, di, dj, thr):
    say(f"cluster 4096x33 nb={nb} di={di} dj={dj} thr={thr}")
    host, dev = matrices(4096, 33, nb, seed=nb * 100 + di * 10 + dj)
    hijs, hits = F.findhits(dev if nb > 1 else dev[0], thr, di=di, dj=dj)
    assert np.array_equal(hijs, H.findprotohits(host if nb > 1 else host[0], thr))
    same(hits, CO.clusterprotohits(host, hijs, di, dj))
    # the same hits from the host-side proto-hit list
    same(F.clusterprotohits(dev if nb > 1 else dev[0], hijs, di, dj), CO.clusterprotohits(host, hijs, di, dj))


def test_repeatable(F):
    host, dev = matrices(4096, 33, 2, seed=5)
    a = F.findhits(dev, 0.5)[1]
    for _ in range(3):
        b = F.findhits(dev, 0.5)[1]
        for x, y in zip(a, b):
            assert np.array_equal(x, y)


def test_one_hit_holds_every_cell(F):
    host, dev = matrices(4096, 33, 2, seed=9)
    hijs, hits = F.findhits(dev, -1.0)
    assert len(hijs) == 4096 * 66 and len(hits.size) == 1
    assert hits.size.tolist() == [4096 * 66] and hits.box.tolist() == [[1, 4096, 1, 66]]
    full = np.concatenate(host, axis=1)
    i, J = np.unravel_index(np.argmax(full.T), full.T.shape)[::-1]
    assert hits.peak.tolist() == [[i + 1, J + 1]] and hits.value[0] == full.max()
    assert np.all(hits.labels == 1)


def test_edges(F):
    host, dev = matrices(512, 8, 1, seed=3, blobs=0)
    hijs, hits = F.findhits(dev[0], 1e9)
    assert len(hijs) == 0 and len(hits.size) == 0 and hits.box.shape == (0, 4) and len(hits.labels) == 0
    one = np.array([[17, 3]])
    same(F.clusterprotohits(dev[0], one), CO.clusterprotohits(host, one))
    # (Nf, 2) and (1, 3) are adjacent linear indices, not neighbours, and
    # neither are (Nf, 5) and (1, 6); (1, 1), (2, 2) and (1, 3) are one hit
    cells = np.array([[1, 1], [2, 2], [512, 2], [1, 3], [512, 5], [1, 6], [512, 8]])
    h = F.clusterprotohits(dev[0], cells)
    same(h, CO.clusterprotohits(host, cells))
    assert h.labels.tolist() == [1, 1, 2, 1, 3, 4, 5]


def test_abi_errors(F):
    import torch
    from fdrt_hip import _lib
    L = _lib.lib()
    host, dev = matrices(4096, 33, 1, seed=11)
    d = dev[0]
    hijs, hits = F.findhits(d, 2.5)
    Hn = len(hits.size)
    assert Hn > 8
    lin = torch.from_numpy((hijs[:, 0] - 1) + 4096 * (hijs[:, 1] - 1)).cuda()
    arr = (C.c_void_p * 1)(d.data_ptr())
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    cap = 5
    peak = torch.zeros(cap, dtype=torch.int64, device="cuda")
    val = torch.zeros(cap, dtype=torch.float32, device="cuda")
    size = torch.zeros(cap, dtype=torch.int64, device="cuda")
    box = torch.zeros(4 * cap, dtype=torch.int64, device="cuda")
    big = 20000  # above the one-workgroup path's 16384 proto-hits
    lab = torch.zeros(max(len(lin), big), dtype=torch.int64, device="cuda")
    n = C.c_int64()

    def call(hp, cnt, capn, di=1, dj=1):
        return L.zdt_clusterprotohits(hp, cnt, arr, 1, 4096, 33, d.stride(1), di, dj, peak.data_ptr(),
                                      val.data_ptr(), size.data_ptr(), box.data_ptr(), capn, C.byref(n),
                                      lab.data_ptr(), st)

    assert call(lin.data_ptr(), len(lin), cap) == _lib.ZDT_ETRUNC and n.value == Hn
    assert np.array_equal(size.cpu().numpy(), hits.size[:cap])
    assert np.array_equal(box.cpu().numpy().reshape(cap, 4) + 1, hits.box[:cap])
    assert np.array_equal(lab[:len(lin)].cpu().numpy() + 1, hits.labels)
    # count only
    assert call(lin.data_ptr(), len(lin), 0) == _lib.ZDT_ETRUNC and n.value == Hn
    assert call(lin.data_ptr(), 0, cap) == _lib.ZDT_OK and n.value == 0
    # out of order, out of range, negative radius
    rev = lin.flip(0).contiguous()
    assert call(rev.data_ptr(), len(rev), cap) == _lib.ZDT_EINVAL
    bad = lin.clone()
    bad[-1] = 4096 * 33
    assert call(bad.data_ptr(), len(bad), cap) == _lib.ZDT_EINVAL
    # out of range first, on the one-workgroup path (this list) and on the
    # general kernels (a list of 20000); on the general kernels that
    # proto-hit is cluster 0, within cap, with no value to take a peak from,
    # and its slot reads peak -1
    wide = torch.arange(0, 3 * big, 3, dtype=torch.int64, device="cuda")
    for b in (bad, wide):
        b[-1] = b[-2] + 1
        for first in (-1, 4096 * 33 + 5):  # negative; past the end and out of order
            b[0] = first
            assert call(b.data_ptr(), len(b), cap) == _lib.ZDT_EINVAL
            if b is wide:
                assert peak[0].item() == -1 and size[0].item() == 0
    # the general kernels on valid input
    wide[0] = 0
    assert call(wide.data_ptr(), big, 0) == _lib.ZDT_ETRUNC and n.value > 0
    assert call(lin.data_ptr(), len(lin), cap, di=-1) == _lib.ZDT_EINVAL
    torch.cuda.synchronize()


# ------------------------------------------------------------------ full size


@pytest.fixture(scope="module")
def cfg2(F):
    say("cluster cfg2 2^20 x 16 x 129")
    ws = F.ZDTWorkspace(synth.cfg2(seed=1234), -4.0, 0.0625, 129)
    fdr = F.zdtfdr(ws)
    return fdr, fdr.cpu().numpy()


def test_full_size_snr8_vs_oracle(F, cfg2):
    fdr, host = cfg2
    hijs, hits = F.findhits(fdr, 8, snr=True)
    say(f"cfg2 SNR 8: {len(hijs)} proto-hits, {len(hits.size)} hits")
    assert np.array_equal(hijs, H.findprotohits(host, F.fdrdenormalize(8, fdr)))
    same(hits, CO.clusterprotohits(host, hijs))
    # against the FP64 oracle's proto-hits (tests/golden/cfg2_hits_snr8.json):
    # 10881 proto-hits in 1401 hits.  The two proto-hit sets differ only in
    # the tie band; one cell more or less changes the count by at most 3 (its
    # 8 neighbours fall into at most 4 groups that it joins or separates)
    gold = json.loads((Path(__file__).parent / "golden" / "cfg2_hits_snr8.json").read_text())
    h64 = CO.components(np.array(gold["hits"]), 1, 1).max() + 1
    assert h64 == 1401 and abs(len(hits.size) - h64) <= 3 * len(gold["band"])


@pytest.mark.parametrize("snr", [3.0, 0.0])
def test_full_size_dense_vs_scipy(F, cfg2, snr):
    from scipy import ndimage
    fdr, host = cfg2
    thr = F.fdrdenormalize(snr, fdr)
    hijs, hits = F.findhits(fdr, thr)
    Hn = len(hits.size)
    say(f"cfg2 threshold {thr:.3f}: {len(hijs)} proto-hits, {Hn} hits")
    # scipy labels the transposed mask so its raster order is column-major,
    # which numbers components by first cell as the hits are numbered
    # (in float64: numpy would round a Python-float threshold to float32)
    maskT = host.T.astype(np.float64) >= thr
    Jm, im = np.nonzero(maskT)
    assert np.array_equal(hijs[:, 0] - 1, im) and np.array_equal(hijs[:, 1] - 1, Jm)
    del Jm, im
    ref, nref = ndimage.label(maskT, structure=np.ones((3, 3), dtype=int))
    assert nref == Hn
    J, i = hijs[:, 1] - 1, hijs[:, 0] - 1
    assert np.array_equal(ref[J, i], hits.labels)
    idx = np.arange(1, Hn + 1)
    assert np.array_equal(np.bincount(hits.labels, minlength=Hn + 1)[1:], hits.size)
    objs = ndimage.find_objects(ref)
    box = np.array([[o[1].start + 1, o[1].stop, o[0].start + 1, o[0].stop] for o in objs])
    assert np.array_equal(box, hits.box)
    assert np.array_equal(ndimage.maximum(host.T, ref, idx).astype(np.float32), hits.value)
    # each peak is a cell of its own hit holding that value
    pj, pi = hits.peak[:, 1] - 1, hits.peak[:, 0] - 1
    assert np.array_equal(ref[pj, pi], idx) and np.array_equal(host[pi, pj], hits.value)
