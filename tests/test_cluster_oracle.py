"""Pins the proto-hit clustering restatement (oracle/cluster_oracle.py) with
checks it does not share code with: scipy.ndimage.label (8-connectivity, the
di = dj = 1 rule), breadth-first search over the brute-force neighbour
relation (any radii), and closed-form cases.  CPU only."""
import numpy as np
import pytest
from scipy import ndimage

from oracle import cluster_oracle as CO


def canon(lab):
    """Labels renumbered by first occurrence: equal arrays <=> equal partitions
    numbered the same way."""
    seen = {}
    return np.array([seen.setdefault(int(x), len(seen)) for x in lab])


def hijs_of(mask):
    """1-based (channel, column) of the True cells, column-major order."""
    J, i = np.nonzero(mask.T)
    return np.stack([i + 1, J + 1], axis=1)


def brute(hijs, di, dj):
    P = len(hijs)
    lab = -np.ones(P, dtype=np.int64)
    nxt = 0
    for s in range(P):
        if lab[s] >= 0:
            continue
        lab[s] = nxt
        todo = [s]
        while todo:
            a = todo.pop()
            for b in range(P):
                if lab[b] < 0 and abs(hijs[a, 0] - hijs[b, 0]) <= di and abs(hijs[a, 1] - hijs[b, 1]) <= dj:
                    lab[b] = nxt
                    todo.append(b)
        nxt += 1
    return lab


@pytest.mark.parametrize("density", [0.05, 0.2, 0.45, 0.6])
def test_eight_connectivity_is_scipy_label(density):
    rng = np.random.default_rng(int(density * 100))
    mask = rng.random((300, 40)) < density
    h = hijs_of(mask)
    lab = CO.components(h, 1, 1)
    ref, nref = ndimage.label(mask, structure=np.ones((3, 3), dtype=int))
    assert lab.max() + 1 == nref
    assert np.array_equal(canon(lab), canon(ref[h[:, 0] - 1, h[:, 1] - 1]))
    # numbered by first member: labels first appear in increasing order
    assert np.array_equal(lab, canon(lab))


@pytest.mark.parametrize("di,dj", [(0, 0), (0, 1), (1, 0), (2, 1), (1, 3), (3, 0), (0, 2), (4, 4)])
def test_any_radius_is_brute_force(di, dj):
    rng = np.random.default_rng(7 * di + dj)
    mask = rng.random((40, 12)) < 0.12
    h = hijs_of(mask)
    assert np.array_equal(CO.components(h, di, dj), brute(h, di, dj))


def test_closed_forms():
    Nf, Nr = 50, 8
    fdr = np.zeros((Nf, Nr), np.float32)
    # filled rectangle: channels 10..14, columns 2..4
    m = np.zeros((Nf, Nr), bool)
    m[9:14, 1:4] = True
    fdr[11, 2] = 5.0
    r = CO.clusterprotohits(fdr, hijs_of(m))
    assert r["size"].tolist() == [15] and r["box"].tolist() == [[10, 14, 2, 4]]
    assert r["peak"].tolist() == [[12, 3]] and r["value"].tolist() == [5.0]
    # diagonal line: one hit when diagonal neighbours count, else one per cell
    d = np.zeros((Nf, Nr), bool)
    for t in range(6):
        d[20 + t, 1 + t] = True
    assert CO.components(hijs_of(d), 1, 1).max() == 0
    assert CO.components(hijs_of(d), 1, 0).tolist() == list(range(6))
    assert CO.components(hijs_of(d), 0, 1).tolist() == list(range(6))
    # adjacent linear indices across a column boundary are not neighbours
    w = np.zeros((Nf, Nr), bool)
    w[Nf - 1, 2] = True
    w[0, 3] = True
    assert CO.components(hijs_of(w), 1, 1).tolist() == [0, 1]


def test_peak_ties_and_batches():
    Nf, Nr = 30, 4
    a = np.zeros((Nf, Nr), np.float32)
    b = np.zeros((Nf, Nr), np.float32)
    # one hit spanning the batch boundary (hcat columns 4 and 5), equal maxima
    m = np.zeros((Nf, 2 * Nr), bool)
    m[5:8, 3:5] = True
    a[6, 3] = 2.0
    b[5, 0] = 2.0
    r = CO.clusterprotohits([a, b], hijs_of(m))
    assert r["size"].tolist() == [6]
    # the tie goes to the earlier proto-hit in column-major order: (7, 4) precedes (6, 5)
    assert r["peak"].tolist() == [[7, 4]] and r["value"].tolist() == [2.0]
    assert r["box"].tolist() == [[6, 8, 4, 5]]


def test_empty():
    r = CO.clusterprotohits(np.zeros((10, 3), np.float32), np.zeros((0, 2), np.int64))
    assert len(r["size"]) == 0 and len(r["labels"]) == 0 and r["box"].shape == (0, 4)
