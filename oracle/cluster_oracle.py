"""Proto-hit clustering restated on the host (test infrastructure only;
nothing in the product imports it).

The reference stops at proto-hits: findprotohits (src/fdrutils.jl:131-169)
returns every FDR cell at or above a threshold, and its docstring
(src/fdrutils.jl:141-144) and README.md:48-53 leave the "clustering of
neighboring proto-hits" into hits to the next stage of a search, outside the
package.  The rule restated here: proto-hits (i, J) and (i', J') are
neighbours when |i - i'| <= di and |J - J'| <= dj, with J the column of the
hcat'd matrices; a hit is a connected component of that graph.  For each hit:
its peak (largest value, ties to the earliest proto-hit in column-major
order), that value, its size and its bounding box; hits are numbered by their
first proto-hit.

The method is deliberately not the device's: cells go into a dict, each
proto-hit looks up its (2di+1)(2dj+1)-1 neighbour cells there, and a
sequential union-find by size merges them.

What pins it (tests/test_cluster_oracle.py):
  * di = dj = 1 is 8-connectivity: the partition equals scipy.ndimage.label
    with a 3x3 structuring element on the dense mask;
  * any radii: the partition equals breadth-first search over the brute-force
    O(P^2) neighbour relation on small inputs;
  * closed forms: a filled rectangle is one hit with that box; a diagonal line
    is one hit under dj >= 1 and one hit per cell under dj = 0; the cells
    (Nf-1, J) and (0, J+1), adjacent in linear index, are two hits.
"""
from __future__ import annotations

from typing import Dict, List, Sequence

import numpy as np


def _values(fdrs: Sequence[np.ndarray], hijs: np.ndarray) -> np.ndarray:
    Nr = fdrs[0].shape[1]
    i = hijs[:, 0] - 1
    J = hijs[:, 1] - 1
    out = np.empty(len(hijs), dtype=np.float32)
    for b, f in enumerate(fdrs):
        sel = (J // Nr) == b
        out[sel] = f[i[sel], J[sel] - b * Nr]
    return out


def components(hijs: np.ndarray, di: int, dj: int) -> np.ndarray:
    """Component labels (0-based, numbered by first member) of 1-based
    proto-hit cells in column-major order."""
    P = len(hijs)
    parent = list(range(P))
    size = [1] * P

    def find(x: int) -> int:
        root = x
        while parent[root] != root:
            root = parent[root]
        while parent[x] != root:
            parent[x], x = root, parent[x]
        return root

    cell: Dict[tuple, int] = {(int(a), int(b)): k for k, (a, b) in enumerate(hijs)}
    for k, (a, b) in enumerate(hijs):
        a, b = int(a), int(b)
        for dJ in range(0, dj + 1):
            for dI in range(-di, di + 1):
                if dJ == 0 and dI <= 0:
                    continue  # each pair once: forward half of the neighbourhood
                m = cell.get((a + dI, b + dJ))
                if m is None:
                    continue
                ra, rb = find(k), find(m)
                if ra != rb:
                    if size[ra] < size[rb]:
                        ra, rb = rb, ra
                    parent[rb] = ra
                    size[ra] += size[rb]
    roots = np.array([find(k) for k in range(P)], dtype=np.int64)
    # number components by their first member
    first: Dict[int, int] = {}
    lab = np.empty(P, dtype=np.int64)
    for k in range(P):
        r = int(roots[k])
        if r not in first:
            first[r] = len(first)
        lab[k] = first[r]
    return lab


def clusterprotohits(fdrs, hijs, di: int = 1, dj: int = 1) -> Dict[str, np.ndarray]:
    """Hits of the 1-based proto-hits ``hijs`` ((P, 2), column-major order) of
    the hcat'd host matrices ``fdrs``.  Returns 1-based peak (H, 2), value
    (H,), size (H,), box (H, 4) = first/last channel, first/last column, and
    labels (P,) = 1-based hit number of each proto-hit."""
    lst: List[np.ndarray] = [fdrs] if isinstance(fdrs, np.ndarray) else list(fdrs)
    h = np.asarray(hijs, dtype=np.int64).reshape(-1, 2)
    lab = components(h, di, dj)
    H = int(lab.max()) + 1 if len(lab) else 0
    v = _values(lst, h)
    size = np.bincount(lab, minlength=H).astype(np.int64)
    box = np.zeros((H, 4), dtype=np.int64)
    box[:, 0] = np.iinfo(np.int64).max
    np.minimum.at(box[:, 0], lab, h[:, 0])
    np.maximum.at(box[:, 1], lab, h[:, 0])
    box[:, 2] = np.iinfo(np.int64).max
    np.minimum.at(box[:, 2], lab, h[:, 1])
    np.maximum.at(box[:, 3], lab, h[:, 1])
    # peak: order by (hit, value descending, proto-hit ordinal ascending)
    order = np.lexsort((np.arange(len(h)), -v.astype(np.float64), lab))
    firsts = order[np.searchsorted(lab[order], np.arange(H))]
    return {"peak": h[firsts], "value": v[firsts], "size": size, "box": box, "labels": lab + 1}
