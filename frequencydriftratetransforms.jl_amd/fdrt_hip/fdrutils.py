"""FDR statistics, normalization and proto-hits on the device (src/fdrutils.jl)."""
from __future__ import annotations

import ctypes as C
from numbers import Number
from typing import List, NamedTuple, Sequence, Tuple, Union

import numpy as np
import torch

from . import _device as D
from . import _lib

Fdrs = Union[torch.Tensor, Sequence[torch.Tensor]]


def create_fdr(spectrogram, rates_or_Nr) -> torch.Tensor:
    """Uninitialized (Nf, Nr) float32 matrix on the spectrogram's device
    (src/fdrutils.jl:11-26)."""
    Nr = rates_or_Nr if isinstance(rates_or_Nr, (int, np.integer)) else len(rates_or_Nr)
    Nf = int(spectrogram.shape[0])
    dev = spectrogram.device if isinstance(spectrogram, torch.Tensor) and spectrogram.is_cuda else D.device()
    return D.empty_colmajor(Nf, int(Nr), torch.float32, dev)


def _as_list(fdrs: Fdrs) -> Tuple[List[torch.Tensor], bool]:
    if isinstance(fdrs, torch.Tensor):
        return [fdrs], True
    return list(fdrs), False


def _ptrs(lst: List[torch.Tensor]):
    Nf, Nr = int(lst[0].shape[0]), int(lst[0].shape[1])
    for f in lst:
        if not isinstance(f, torch.Tensor) or not f.is_cuda or f.dtype != torch.float32:
            raise TypeError("FDR matrices must be float32 device tensors")
        if tuple(f.shape) != (Nf, Nr):
            raise ValueError("all FDR matrices must have the same size")
        if f.stride(0) != 1:
            raise ValueError("FDR matrices must be column-major")
    ld = int(lst[0].stride(1)) if Nr > 1 else Nf
    for f in lst:
        if Nr > 1 and f.stride(1) != ld:
            raise ValueError("all FDR matrices must have the same column stride")
    arr = (C.c_void_p * len(lst))(*[f.data_ptr() for f in lst])
    return arr, Nf, Nr, ld


def fdrstats(fdrs: Fdrs) -> Tuple[float, float]:
    """(mean of the first column of the first matrix, minimum over all columns
    of the corrected standard deviation)  (src/fdrutils.jl:39-50)."""
    lst, _ = _as_list(fdrs)
    arr, Nf, Nr, ld = _ptrs(lst)
    m = C.c_double()
    s = C.c_double()
    _lib.check(_lib.lib().zdt_fdrstats(arr, len(lst), Nf, Nr, ld, C.c_void_p(D.stream_ptr()), C.byref(m),
                                       C.byref(s)))
    return float(m.value), float(s.value)


def fdrnormalize(x, m, s):
    """(x - m)/s for a number (src/fdrutils.jl:65-68); for matrices returns
    normalized copies."""
    if isinstance(x, Number):
        return (x - m) / s
    lst, single = _as_list(x)
    copies = [f.clone() for f in lst]
    fdrnormalize_(copies, m, s)
    return copies[0] if single else copies


def fdrnormalize_(fdrs: Fdrs, m=None, s=None):
    """x .= (x .- m) ./ s in place, stats from fdrstats when not given
    (src/fdrutils.jl:70-90)."""
    lst, single = _as_list(fdrs)
    if m is None or s is None:
        m, s = fdrstats(lst)
    arr, Nf, Nr, ld = _ptrs(lst)
    _lib.check(_lib.lib().zdt_fdrnormalize(arr, len(lst), Nf, Nr, ld, float(m), float(s),
                                           C.c_void_p(D.stream_ptr())))
    return fdrs


def fdrdenormalize(snr, *args) -> float:
    """snr*s + m, with (m, s) given or computed from fdr(s)
    (src/fdrutils.jl:104-119)."""
    if len(args) == 2:
        m, s = args
    elif len(args) == 1:
        m, s = fdrstats(args[0])
    else:
        raise TypeError("fdrdenormalize(snr, m, s) or fdrdenormalize(snr, fdr[s])")
    return float(snr) * float(s) + float(m)


def _protohits_device(lst: List[torch.Tensor], threshold: float, capacity: int):
    """zdt_findprotohits into a device buffer: (buffer, count, Nf, Nr, ld, ptrs)."""
    arr, Nf, Nr, ld = _ptrs(lst)
    L = _lib.lib()
    n = C.c_int64()
    st = C.c_void_p(D.stream_ptr())
    cap = int(capacity)
    while True:
        buf = torch.empty(max(cap, 1), dtype=torch.int64, device=lst[0].device)
        rc = L.zdt_findprotohits(arr, len(lst), Nf, Nr, ld, float(threshold), buf.data_ptr(), cap, C.byref(n), st)
        if rc == _lib.ZDT_ETRUNC:
            cap = int(n.value)
            continue
        _lib.check(rc)
        return buf, int(n.value), Nf, Nr, ld, arr


def findprotohits(fdrs: Fdrs, threshold, snr: bool = False, capacity: int = 1 << 20) -> np.ndarray:
    """Every element >= threshold as 1-based (channel, rate) pairs, column-major
    order over the hcat'd matrices (src/fdrutils.jl:148-169).  Returns an
    (H, 2) int64 array, the counterpart of Vector{CartesianIndex{2}}."""
    lst, _ = _as_list(fdrs)
    if snr:
        threshold = fdrdenormalize(threshold, lst)
    buf, n, Nf, _, _, _ = _protohits_device(lst, threshold, capacity)
    lin = buf[:n].cpu().numpy()
    i = lin % Nf
    j = lin // Nf
    return np.stack([i + 1, j + 1], axis=1).astype(np.int64)


class Hits(NamedTuple):
    """Hits: connected groups of neighbouring proto-hits, numbered in the
    order of their first proto-hit (column-major).  Coordinates are 1-based
    (channel, hcat'd rate column), like findprotohits' proto-hits."""
    peak: np.ndarray    # (H, 2) the cell of the largest value (ties: the earliest proto-hit)
    value: np.ndarray   # (H,) float32 that value
    size: np.ndarray    # (H,) number of proto-hits
    box: np.ndarray     # (H, 4) first channel, last channel, first column, last column
    labels: np.ndarray  # (P,) 1-based hit number of each proto-hit


def _cluster_device(buf: torch.Tensor, n: int, arr, nb: int, Nf: int, Nr: int, ld: int, di: int, dj: int,
                    capacity: int) -> Hits:
    L = _lib.lib()
    st = C.c_void_p(D.stream_ptr())
    dev = buf.device
    label = torch.empty(max(n, 1), dtype=torch.int64, device=dev)
    ncl = C.c_int64()
    cap = int(capacity)
    while True:
        c1 = max(cap, 1)
        peak = torch.empty(c1, dtype=torch.int64, device=dev)
        value = torch.empty(c1, dtype=torch.float32, device=dev)
        size = torch.empty(c1, dtype=torch.int64, device=dev)
        box = torch.empty(4 * c1, dtype=torch.int64, device=dev)
        rc = L.zdt_clusterprotohits(buf.data_ptr(), n, arr, nb, Nf, Nr, ld, int(di), int(dj), peak.data_ptr(),
                                    value.data_ptr(), size.data_ptr(), box.data_ptr(), cap, C.byref(ncl),
                                    label.data_ptr(), st)
        if rc == _lib.ZDT_ETRUNC:
            cap = int(ncl.value)
            continue
        _lib.check(rc)
        break
    h = int(ncl.value)
    pk = peak[:h].cpu().numpy()
    bx = box[: 4 * h].cpu().numpy().reshape(h, 4)
    return Hits(peak=np.stack([pk % Nf + 1, pk // Nf + 1], axis=1).astype(np.int64),
                value=value[:h].cpu().numpy(), size=size[:h].cpu().numpy(), box=bx + 1,
                labels=label[:n].cpu().numpy() + 1)


def clusterprotohits(fdrs: Fdrs, hijs, di: int = 1, dj: int = 1, capacity: int = 1 << 16) -> Hits:
    """Group the proto-hits ``hijs`` ((P, 2) 1-based channel/column pairs in
    column-major order, as findprotohits returns them) into hits: two
    proto-hits are neighbours when their channels differ by at most ``di`` and
    their columns by at most ``dj``, and a hit is a connected group.  The
    reference leaves this step to the stage after findprotohits
    (src/fdrutils.jl:141-144, README.md:48-53)."""
    lst, _ = _as_list(fdrs)
    arr, Nf, Nr, ld = _ptrs(lst)
    h = np.asarray(hijs, dtype=np.int64).reshape(-1, 2)
    lin = (h[:, 0] - 1) + Nf * (h[:, 1] - 1)
    buf = torch.from_numpy(np.ascontiguousarray(lin)).to(lst[0].device) if len(lin) else \
        torch.empty(1, dtype=torch.int64, device=lst[0].device)
    return _cluster_device(buf, len(lin), arr, len(lst), Nf, Nr, ld, di, dj, capacity)


def findhits(fdrs: Fdrs, threshold, snr: bool = False, di: int = 1, dj: int = 1,
             capacity: int = 1 << 20) -> Tuple[np.ndarray, Hits]:
    """findprotohits followed by clusterprotohits, with the proto-hits kept
    on the device between the two.  Returns (proto-hits, hits)."""
    lst, _ = _as_list(fdrs)
    if snr:
        threshold = fdrdenormalize(threshold, lst)
    buf, n, Nf, Nr, ld, arr = _protohits_device(lst, threshold, capacity)
    hits = _cluster_device(buf, n, arr, len(lst), Nf, Nr, ld, di, dj, 1 << 16)
    lin = buf[:n].cpu().numpy()
    return np.stack([lin % Nf + 1, lin // Nf + 1], axis=1).astype(np.int64), hits
