"""Long-integration search (include/zdt.h, zdt_stream_*): spectrograms of one
shape streamed from host memory, rates cut into equal HBM-sized batches, and
proto-hits at an SNR over all batches of each spectrogram.

It replaces the loop a reference user writes around ``batchrates``
(src/batchrates.jl:22-60), ``zdtfdr!(dests, ws; r0=...)``
(src/zdtfdr.jl:349-362) and ``findprotohits(fdrs, snr; snr=true)``
(src/fdrutils.jl:148-169).
"""
from __future__ import annotations

import ctypes as C
from typing import List, Sequence, Tuple

import numpy as np
import torch

from . import _lib
from .zdtutils import float32_of, rates_params


def pinned_spectrogram(Nf: int, Nt: int) -> torch.Tensor:
    """An (Nf, Nt) column-major float32 host tensor in pinned memory, whose
    copies to the device overlap the search."""
    return torch.empty((Nt, Nf), dtype=torch.float32, pin_memory=True).t()


class StreamSearch:
    """StreamSearch(Nf, Nt, rates; Nrb=0, hbm_bytes=0, generic=False, unnormalized=False)

    ``rates``: a range ``ZDTWorkspace`` accepts, or an (r0, dr, Nr) tuple.
    ``Nrb=0`` and ``hbm_bytes=0`` let the library size the batches to 90% of
    the free device memory.  ``search(specs, snr)`` returns one (H, 2) array
    of 1-based (channel, rate-column) indices per spectrogram (columns of the
    hcat'd batches, i.e. rate index + 1) and the thresholds used.  When Nr is
    not a multiple of the batch length the last batch computes rates past
    r0 + (Nr-1)*dr, but only the Nr requested columns enter the statistics
    and the hits.  ``profile=True`` times every kernel of the searches (HIP
    events on the search's stream; ``kernel_times``, ``reset_profile``).
    """

    def __init__(self, Nf: int, Nt: int, rates, *, Nrb: int = 0, hbm_bytes: int = 0, generic: bool = False,
                 unnormalized: bool = False, precision: str = "fp32", profile: bool = False):
        if isinstance(rates, tuple) and len(rates) == 3:
            r0, dr, Nr = float32_of(rates[0]), float32_of(rates[1]), int(rates[2])
        else:
            r0, dr, Nr = rates_params(rates)
        self.Nf, self.Nt, self.Nr = int(Nf), int(Nt), int(Nr)
        self.r0, self.dr = np.float32(r0), np.float32(dr)
        flags = (_lib.ZDT_FLAG_GENERIC if generic else 0) | (_lib.ZDT_FLAG_UNNORMALIZED if unnormalized else 0)
        flags |= _lib.ZDT_FLAG_PROFILE if profile else 0
        if precision not in _lib.PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(_lib.PRECISIONS)}")
        self.precision = precision
        h = C.c_void_p()
        self._h = None
        _lib.check(_lib.lib().zdt_stream_create(self.Nf, self.Nt, float(self.r0), float(self.dr), self.Nr, int(Nrb),
                                                int(hbm_bytes), flags, _lib.PRECISIONS[precision], C.byref(h)))
        self._h = h
        nrb, nb, res, by = C.c_int64(), C.c_int64(), C.c_int(), C.c_int64()
        _lib.check(_lib.lib().zdt_stream_info(self._h, C.byref(nrb), C.byref(nb), C.byref(res), C.byref(by)))
        self.Nrb, self.nbatch, self.resident, self.nbytes = nrb.value, nb.value, bool(res.value), by.value

    def _workspace(self) -> C.c_void_p:
        ws = C.c_void_p()
        _lib.check(_lib.lib().zdt_stream_workspace(self._h, C.byref(ws)))
        return ws

    @property
    def path(self) -> str:
        """The kernel path of the search's workspace (as ZDTWorkspace.path)."""
        p = C.c_int()
        _lib.check(_lib.lib().zdt_path(self._workspace(), C.byref(p), None))
        return _lib.PATHS[p.value]

    def reset_profile(self) -> None:
        _lib.check(_lib.lib().zdt_profile_reset(self._workspace()))

    def kernel_times(self) -> List[Tuple[str, float, int]]:
        """(kernel, total ms, launches) over the searches since the last
        reset_profile (profile=True), longest first."""
        L = _lib.lib()
        ws = self._workspace()
        cnt = C.c_int()
        _lib.check(L.zdt_profile_read(ws, -1, C.byref(cnt), None, 0, None, None, None))
        out = []
        for i in range(cnt.value):
            name = C.create_string_buffer(64)
            ms, nl, by = C.c_double(), C.c_int64(), C.c_double()
            _lib.check(L.zdt_profile_read(ws, i, None, name, 64, C.byref(ms), C.byref(nl), C.byref(by)))
            out.append((name.value.decode(), ms.value, nl.value))
        return sorted(out, key=lambda x: -x[1])

    def close(self) -> None:
        if self._h is not None:
            _lib.lib().zdt_stream_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def search(self, specs: Sequence, snr: float = 10.0, cap: int = 1 << 20) -> Tuple[List[np.ndarray], np.ndarray]:
        keep, ptrs = self._pointers(specs)
        n = len(ptrs)
        hits = np.zeros(max(1, n * cap), dtype=np.int64)
        nh = np.zeros(n, dtype=np.int64)
        thr = np.zeros(n, dtype=np.float64)
        rc = _lib.lib().zdt_stream_search(self._h, (C.c_void_p * n)(*ptrs), n, float(snr),
                                          hits.ctypes.data_as(C.c_void_p), int(cap), nh.ctypes.data_as(C.c_void_p),
                                          thr.ctypes.data_as(C.c_void_p))
        _lib.check(rc)
        out = []
        for i in range(n):
            lin = hits[i * cap: i * cap + nh[i]]
            out.append(np.stack([lin % self.Nf + 1, lin // self.Nf + 1], axis=1).astype(np.int64)
                       if nh[i] else np.zeros((0, 2), np.int64))
        return out, thr

    def findhits(self, specs: Sequence, snr: float = 10.0, di: int = 1, dj: int = 1, cap: int = 1 << 16,
                 pcap: int = 1 << 22) -> Tuple[List["Hits"], np.ndarray, np.ndarray]:
        """The search, ending in hits: each spectrogram's proto-hits are
        clustered on the device (as ``fdrt_hip.clusterprotohits`` clusters
        them) while its batches' FDRs are resident.  Returns (one ``Hits`` per
        spectrogram, with empty ``labels``; the proto-hit counts; the
        thresholds).  Needs a resident stream (``self.resident``)."""
        from .fdrutils import Hits
        keep, ptrs = self._pointers(specs)
        n = len(ptrs)
        c1 = max(int(cap), 1)
        peak = np.zeros(n * c1, np.int64)
        value = np.zeros(n * c1, np.float32)
        size = np.zeros(n * c1, np.int64)
        box = np.zeros(4 * n * c1, np.int64)
        ncl = np.zeros(n, np.int64)
        nph = np.zeros(n, np.int64)
        thr = np.zeros(n, np.float64)
        vp = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
        rc = _lib.lib().zdt_stream_findhits(self._h, (C.c_void_p * n)(*ptrs), n, float(snr), int(di), int(dj),
                                            int(pcap), vp(peak), vp(value), vp(size), vp(box), int(cap), vp(ncl),
                                            vp(nph), vp(thr))
        _lib.check(rc)
        out = []
        for i in range(n):
            h = int(ncl[i])
            o = i * int(cap)
            pk = peak[o:o + h]
            out.append(Hits(peak=np.stack([pk % self.Nf + 1, pk // self.Nf + 1], axis=1).astype(np.int64),
                            value=value[o:o + h].copy(), size=size[o:o + h].copy(),
                            box=box[4 * o:4 * (o + h)].reshape(h, 4) + 1, labels=np.zeros(0, np.int64)))
        return out, nph, thr

    def _pointers(self, specs: Sequence):
        keep = []
        ptrs = []
        for s in specs:
            if isinstance(s, torch.Tensor):
                if s.is_cuda or s.dtype != torch.float32 or tuple(s.shape) != (self.Nf, self.Nt):
                    raise ValueError(f"expected host float32 ({self.Nf}, {self.Nt}) spectrograms")
                if not (s.stride(0) == 1 and s.stride(1) == self.Nf):
                    s = s.t().contiguous().t()
                keep.append(s)
                ptrs.append(s.data_ptr())
            else:
                a = np.asfortranarray(np.asarray(s, dtype=np.float32))
                if a.shape != (self.Nf, self.Nt):
                    raise ValueError(f"expected ({self.Nf}, {self.Nt}) spectrograms")
                keep.append(a)
                ptrs.append(a.ctypes.data)
        return keep, ptrs

    def search_files(self, paths: Sequence, snr: float = 10.0, cap: int = 1 << 20, channels=None, ifnum: int = 0,
                     times=None, group: int = 2, dataset: str = "data") -> Tuple[List[np.ndarray], np.ndarray]:
        """search() over filterbank HDF5 files (fdrt_hip.filterbank): each
        file's data[channels, ifnum, times] spectrogram (Nf x Nt) is read
        straight into pinned host memory on a reader thread, ``group`` files
        at a time, while the GPU searches the group before; two groups of
        pinned buffers alternate.  The reads (HDF5, bitshuffle decoding
        included) release the GIL, so they overlap the search.  Returns
        (hits, thresholds) as search() does, one entry per file, in order."""
        from concurrent.futures import ThreadPoolExecutor

        from .filterbank import read_spectrogram
        paths = list(paths)
        if not paths:
            return [], np.zeros(0)
        group = max(1, int(group))
        groups = [paths[i:i + group] for i in range(0, len(paths), group)]
        bufs = [[pinned_spectrogram(self.Nf, self.Nt) for _ in range(len(groups[0]))] for _ in range(min(2, len(groups)))]

        def fill(k, names):
            for buf, path in zip(bufs[k], names):
                read_spectrogram(path, channels=channels, ifnum=ifnum, times=times, out=buf, dataset=dataset)
            return len(names)

        hits, thrs = [], []
        with ThreadPoolExecutor(max_workers=1) as reader:
            pending = reader.submit(fill, 0, groups[0])
            for g in range(len(groups)):
                m = pending.result()
                if g + 1 < len(groups):  # the next group's reads run during this search
                    pending = reader.submit(fill, (g + 1) % 2, groups[g + 1])
                h, t = self.search(bufs[g % 2][:m], snr=snr, cap=cap)
                hits += h
                thrs.append(t)
        return hits, np.concatenate(thrs)
