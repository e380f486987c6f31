"""Ragged multi-channel plans (include/zdt.h, zdt_plan_*): many spectrograms
of uneven Nf and Nt through input!, one zdtfdr! batch, fdrdenormalize and
findprotohits, replayed as one hipGraph launch.

The reference has no such object; it replaces the host loop a reference
user writes over channels (``zdtfdr!`` per channel, src/zdtfdr.jl:349-421,
then ``findprotohits(fdr, snr; snr=true)``, src/fdrutils.jl:148-169).
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np
import torch

from . import _device as D
from . import _lib
from .zdtutils import float32_of, rates_params


class RaggedPlan:
    """RaggedPlan(specs, rates; autotune=True, unnormalized=False, generic=False)

    ``specs``: list of (Nf_c, Nt_c) spectrograms (column-major device tensors or
    numpy arrays); ``rates``: one rate range per channel (anything
    ``ZDTWorkspace`` accepts as rates, or an (r0, dr, Nr) tuple).
    ``bind(snr=..., cap=...)`` captures the graph; ``launch()`` replays it;
    ``hits(c)`` returns channel c's proto-hits as an (H, 2) int64 array of
    1-based (channel row, rate column) indices, like ``findprotohits``.
    """

    def __init__(self, specs: Sequence, rates: Sequence, *, autotune: bool = True, unnormalized: bool = False,
                 generic: bool = False):
        if len(specs) != len(rates) or len(specs) == 0:
            raise ValueError("need one rate range per spectrogram")
        self.specs = [D.colmajor(s) for s in specs]
        params = []
        for r in rates:
            if isinstance(r, tuple) and len(r) == 3:
                params.append((float32_of(r[0]), float32_of(r[1]), int(r[2])))
            else:
                params.append(rates_params(r))
        self.nch = len(self.specs)
        self.Nf = [int(s.shape[0]) for s in self.specs]
        self.Nt = [int(s.shape[1]) for s in self.specs]
        self.r0 = [np.float32(p[0]) for p in params]
        self.dr = [np.float32(p[1]) for p in params]
        self.Nr = [int(p[2]) for p in params]
        flags = ((_lib.ZDT_PLAN_AUTOTUNE if autotune else 0) | (_lib.ZDT_FLAG_UNNORMALIZED if unnormalized else 0)
                 | (_lib.ZDT_FLAG_GENERIC if generic else 0))
        L = _lib.lib()
        i64 = C.c_int64 * self.nch
        f32 = C.c_float * self.nch
        h = C.c_void_p()
        self._h = None
        _lib.check(L.zdt_plan_ragged(self.nch, i64(*self.Nf), i64(*self.Nt), f32(*map(float, self.r0)),
                                     f32(*map(float, self.dr)), i64(*self.Nr), flags, C.c_void_p(D.stream_ptr()),
                                     C.byref(h)))
        self._h = h
        self.dests: Optional[list] = None
        self.cap = 0
        self._hits = None
        self._nhits = None
        self._thr = None

    def close(self) -> None:
        if self._h is not None:
            _lib.lib().zdt_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self, c: int) -> dict:
        """Tuned choices for channel c: shape index, path, n2 (fused), maxr (generic)."""
        v = [C.c_int() for _ in range(4)]
        _lib.check(_lib.lib().zdt_plan_info(self._h, int(c), *[C.byref(x) for x in v]))
        return {"shape": v[0].value, "path": {1: "fused", 2: "fused-output"}.get(v[1].value, "generic"),
                "n2": v[2].value, "maxr": v[3].value}

    def tuning(self, c: int) -> dict:
        """Where channel c's choices came from ("default", "timed" or "table")
        and the chosen candidate's time in ms (input! + one batch)."""
        src = C.c_int()
        ms = C.c_float()
        _lib.check(_lib.lib().zdt_plan_tuning(self._h, int(c), C.byref(src), C.byref(ms)))
        return {"source": {0: "default", 1: "timed", 2: "table"}[src.value], "ms": ms.value}

    def bind(self, thresholds=10.0, *, snr: bool = True, cap: int = 4096, keep_fdrs: bool = True) -> None:
        """Captures the graph.  thresholds: one value or one per channel (SNRs
        when snr=True).  keep_fdrs=False lets channels of one shape share a
        plan-owned FDR buffer (only the hits survive)."""
        thr = np.broadcast_to(np.asarray(thresholds, dtype=np.float64), (self.nch,)).copy()
        dev = self.specs[0].device
        self.dests = [D.empty_colmajor(nf, nr, dev=dev) for nf, nr in zip(self.Nf, self.Nr)] if keep_fdrs else None
        self.cap = int(cap)
        self._hits = torch.empty(max(1, self.nch * self.cap), dtype=torch.int64, device=dev)
        self._nhits = torch.zeros(self.nch, dtype=torch.int64, device=dev)
        self._thr = torch.zeros(self.nch, dtype=torch.float64, device=dev)
        vp = C.c_void_p * self.nch
        specs = vp(*[s.data_ptr() for s in self.specs])
        dests = vp(*[d.data_ptr() for d in self.dests]) if self.dests is not None else None
        torch.cuda.current_stream().synchronize()
        _lib.check(_lib.lib().zdt_plan_bind(self._h, specs, dests, (C.c_double * self.nch)(*thr), int(bool(snr)),
                                            C.c_void_p(self._hits.data_ptr()), self.cap,
                                            C.c_void_p(self._nhits.data_ptr()), C.c_void_p(self._thr.data_ptr())))

    def launch(self) -> None:
        _lib.check(_lib.lib().zdt_plan_launch(self._h))

    def nhits(self) -> np.ndarray:
        return self._nhits.cpu().numpy()

    def thresholds(self) -> np.ndarray:
        return self._thr.cpu().numpy()

    def hits(self, c: int) -> np.ndarray:
        n = int(self._nhits[c].item())
        if n > self.cap:
            raise _lib.ZDTError(_lib.ZDT_ETRUNC, f"channel {c}: {n} hits exceed capacity {self.cap}")
        lin = self._hits[c * self.cap: c * self.cap + n].cpu().numpy()
        Nf = self.Nf[c]
        return np.stack([lin % Nf + 1, lin // Nf + 1], axis=1).astype(np.int64) if n else np.zeros((0, 2), np.int64)

    def fdr(self, c: int) -> torch.Tensor:
        if self.dests is None:
            raise RuntimeError("bind(keep_fdrs=True) keeps the FDRs")
        return self.dests[c]


def set_tune_table(path: Optional[str]) -> None:
    """Names the tuning table autotuned plans read and append to (None turns it
    off; process-wide, zdt_tune_table)."""
    _lib.check(_lib.lib().zdt_tune_table(None if path is None else str(path).encode()))
