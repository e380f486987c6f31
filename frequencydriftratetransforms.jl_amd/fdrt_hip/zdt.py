"""ZDTWorkspace and the ZDT step functions over libzdt_hip.so.

Python mirror of src/zdtfdr.jl (and of the Julia binding julia/ZDTHip, a
package extension of FrequencyDriftRateTransforms): same names with ``!``
spelled as a trailing underscore, same argument meaning, same errors.  Device arrays are torch tensors of the
Julia shape with column-major strides (see _device.py).
"""
from __future__ import annotations

import ctypes as C
import itertools
from typing import Iterable, Optional, Sequence, Union

import numpy as np
import torch

from . import _device as D
from . import _lib
from .zdtutils import RateRange, driftrates, float32_of, rates_params

_WINDOW_NAMES = {"rect": _lib.ZDT_WINDOW_RECT, "hamming": _lib.ZDT_WINDOW_HAMMING}


def _window(w, Nf: int):
    """(kind, device tensor or None) for a window given as "rect"/"hamming",
    a two-argument function w(n, N), or a vector of Nf//2+1 values
    (src/zdtfdr.jl:297-321)."""
    Nk = Nf // 2 + 1
    if w is None:
        return _lib.ZDT_WINDOW_RECT, None
    if isinstance(w, str):
        if w not in _WINDOW_NAMES:
            raise ValueError(f"unsupported window type ({w})")
        return _WINDOW_NAMES[w], None
    if callable(w):
        vals = np.array([float(w(n, Nf)) for n in range(Nk)], dtype=np.float64)
        t = torch.from_numpy(vals.astype(np.float32)).to(D.device())
        return _lib.ZDT_WINDOW_CUSTOM, t
    t = torch.as_tensor(np.asarray(w) if not isinstance(w, torch.Tensor) else w)
    t = t.to(device=D.device(), dtype=torch.float32).contiguous()
    if t.dim() != 1 or t.numel() != Nk:
        raise ValueError(f"window vector must have Nf//2+1 = {Nk} entries")
    return _lib.ZDT_WINDOW_CUSTOM, t


def _ptr(t: Optional[torch.Tensor]) -> Optional[int]:
    return None if t is None else t.data_ptr()


class ZDTWorkspace:
    """Workspace of the Chirp-Z De-Doppler Transform on the GPU.

    ``ZDTWorkspace(spectrogram, rates, factors=(2,3,5))`` or
    ``ZDTWorkspace(spectrogram, r0, dr, Nr, factors=(2,3,5))``
    (src/zdtfdr.jl:37-87).  ``spectrogram`` is an (Nf, Nt) float32 matrix
    (torch tensor or numpy array; copied to the device in column-major order
    if needed), frequency fastest.  Construction computes Nl, allocates the
    device buffers and runs ``input_`` on the spectrogram.

    Keyword-only extras: ``unnormalized=True`` scales the output by Nf*Nl
    (test/runtests.jl:31 convention); ``generic=True`` forces the unfused
    kernels; ``n2=512`` selects the fused path's 512-point strided FFTs
    where 1024 is the default; ``radix_cap=8`` limits the generic FFTs to
    radices up to 8 (16 is the default); ``precision`` is "fp32" (default),
    "split-bf16" or "split-fp16" (the fused path's DFT stages as MFMA
    products of split hi+lo operands).
    """

    def __init__(self, spectrogram, *args, factors=None, output_aligned: bool = False,
                 unnormalized: bool = False, generic: bool = False, n2: int = 0, radix_cap: int = 16,
                 precision: str = "fp32"):
        if len(args) in (1, 2) and not isinstance(args[0], (int, float, np.floating, np.integer)):
            r0, dr, Nr = rates_params(args[0])
            if len(args) == 2:
                factors = args[1]
        elif len(args) in (3, 4):
            r0, dr, Nr = float32_of(args[0]), float32_of(args[1]), int(args[2])
            if len(args) == 4:
                factors = args[3]
        else:
            raise TypeError("ZDTWorkspace(spectrogram, rates[, factors]) or "
                            "ZDTWorkspace(spectrogram, r0, dr, Nr[, factors])")
        self.factors = tuple(int(f) for f in (factors if factors is not None else (2, 3, 5)))
        self.output_aligned = bool(output_aligned)
        spec = D.colmajor(spectrogram)
        self.Nf, self.Nt = int(spec.shape[0]), int(spec.shape[1])
        self.r0 = np.float32(r0)
        self.dr = np.float32(dr)
        self.Nr = int(Nr)
        self.device = spec.device
        self._stream = D.stream_ptr()
        flags = ((_lib.ZDT_FLAG_UNNORMALIZED if unnormalized else 0) | (_lib.ZDT_FLAG_GENERIC if generic else 0)
                 | (_lib.ZDT_FLAG_N2_512 if n2 == 512 else 0) | (_lib.ZDT_FLAG_RADIX8 if radix_cap == 8 else 0))
        if radix_cap not in (8, 16):
            raise ValueError("radix_cap must be 8 or 16")
        if precision not in _lib.PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(_lib.PRECISIONS)}")
        arr, n = _lib.int_array(self.factors)
        h = C.c_void_p()
        self._h = None
        _lib.check(_lib.lib().zdt_create(spec.data_ptr(), self.Nf, self.Nt, float(self.r0), float(self.dr), self.Nr,
                                         arr, n, flags, _lib.PRECISIONS[precision], C.c_void_p(self._stream),
                                         C.byref(h)))
        self._h = h
        nl = C.c_int64()
        _lib.check(_lib.lib().zdt_info(self._h, None, None, None, None, None, C.byref(nl)))
        self.Nl = int(nl.value)
        self._window_cache = {}

    # -- lifetime ---------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "_h", None) is not None:
            _lib.lib().zdt_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self) -> C.c_void_p:
        if self._h is None:
            raise RuntimeError("workspace was closed")
        return self._h

    # -- reference fields ---------------------------------------------------
    @property
    def δr(self) -> np.float32:  # noqa: N802 (reference field name)
        return self.dr

    @property
    def Nk(self) -> int:
        return self.Nf // 2 + 1

    @property
    def path(self) -> str:
        """Kernel path the library runs for this workspace: "fused",
        "fused-output" or "generic" (zdt_path)."""
        v = C.c_int()
        _lib.check(_lib.lib().zdt_path(self.handle, C.byref(v), None))
        return _lib.PATHS[v.value]

    @property
    def precision(self) -> str:
        """Arithmetic of the DFT stages: "fp32", "split-bf16" or "split-fp16"."""
        v = C.c_int()
        _lib.check(_lib.lib().zdt_path(self.handle, None, C.byref(v)))
        return {b: a for a, b in _lib.PRECISIONS.items()}[v.value]

    def _buffer(self, which: int) -> torch.Tensor:
        p = C.c_void_p()
        rows = C.c_int64()
        cols = C.c_int64()
        _lib.check(_lib.lib().zdt_buffer(self.handle, which, C.byref(p), C.byref(rows), C.byref(cols)))
        return D.copy_from_ptr(p.value, rows.value, cols.value, True)

    @property
    def F(self) -> torch.Tensor:
        """Copy of the input spectrum F (Nk x Nt complex64)."""
        return self._buffer(_lib.ZDT_BUF_F)

    @property
    def Y(self) -> torch.Tensor:
        """Copy of the CZT work matrix Y (Nk x Nl) of the step functions."""
        return self._buffer(_lib.ZDT_BUF_Y)

    @property
    def Ys(self) -> torch.Tensor:
        """Copy of the first Nr columns of Y."""
        return self.Y[:, : self.Nr]

    @property
    def V(self) -> torch.Tensor:
        """Copy of the chirp-kernel spectrum V (Nk x Nl)."""
        return self._buffer(_lib.ZDT_BUF_V)

    @property
    def nbytes(self) -> int:
        """Device bytes held (Base.sizeof, src/zdtfdr.jl:101-103)."""
        b = C.c_int64()
        _lib.check(_lib.lib().zdt_sizeof(self.handle, C.byref(b)))
        return int(b.value)

    def __repr__(self) -> str:
        return (f"ZDTWorkspace(Nf={self.Nf},Nt={self.Nt},r0={float(self.r0)},δr={float(self.dr)},"
                f"Nr={self.Nr},Nl={self.Nl},{self.factors})")

    def window(self, w):
        """(kind, device values) of a window; named windows are cached."""
        if w is None or isinstance(w, str):
            if w not in self._window_cache:
                self._window_cache[w] = _window(w, self.Nf)
            return self._window_cache[w]
        return _window(w, self.Nf)

    def set_stream(self, stream_ptr: int) -> None:
        self._stream = stream_ptr
        _lib.check(_lib.lib().zdt_set_stream(self.handle, C.c_void_p(stream_ptr)))


# ---------------------------------------------------------------------------
# step functions  (src/zdtfdr.jl:209-331)
# ---------------------------------------------------------------------------


def _check_spec(ws: ZDTWorkspace, spectrogram) -> torch.Tensor:
    spec = D.colmajor(spectrogram)
    if tuple(spec.shape) != (ws.Nf, ws.Nt):
        raise ValueError(f"spectrogram has size {tuple(spec.shape)}, workspace expects {(ws.Nf, ws.Nt)}")
    return spec


def input_(ws: ZDTWorkspace, spectrogram) -> None:
    """input!(workspace, spectrogram): F = rfft(spectrogram)  (src/zdtfdr.jl:209-212)."""
    spec = _check_spec(ws, spectrogram)
    _lib.check(_lib.lib().zdt_input(ws.handle, spec.data_ptr()))


def preprocess_(ws: ZDTWorkspace, r0=None) -> None:
    """preprocess!(workspace, r0=workspace.r0)  (src/zdtfdr.jl:221-239)."""
    r0 = ws.r0 if r0 is None else float32_of(r0)
    _lib.check(_lib.lib().zdt_preprocess(ws.handle, float(r0)))


def convolve_(ws: ZDTWorkspace) -> None:
    """convolve!(workspace): fft along dim 2, times V, inverse fft (src/zdtfdr.jl:249-253)."""
    _lib.check(_lib.lib().zdt_convolve(ws.handle))


def postprocess_(*args, w=None) -> None:
    """postprocess!([w,] workspace)  (src/zdtfdr.jl:297-321)."""
    if len(args) == 2:
        w, ws = args
    elif len(args) == 1:
        ws = args[0]
    else:
        raise TypeError("postprocess_([w,] workspace)")
    kind, wt = ws.window(w)
    _lib.check(_lib.lib().zdt_postprocess(ws.handle, kind, _ptr(wt)))


def _check_dest(ws: ZDTWorkspace, dest: torch.Tensor) -> torch.Tensor:
    if not isinstance(dest, torch.Tensor) or not dest.is_cuda or dest.dtype != torch.float32:
        raise TypeError("dest must be a float32 device tensor")
    if tuple(dest.shape) != (ws.Nf, ws.Nr) or not D.is_colmajor(dest):
        raise ValueError(f"dest must be a column-major {(ws.Nf, ws.Nr)} matrix")
    return dest


def output_(dest: torch.Tensor, ws: ZDTWorkspace) -> torch.Tensor:
    """output!(dest, workspace): inverse real FFT of Ys into dest (src/zdtfdr.jl:328-331)."""
    _check_dest(ws, dest)
    _lib.check(_lib.lib().zdt_output(ws.handle, dest.data_ptr()))
    return dest


def create_fdr_for(ws: ZDTWorkspace) -> torch.Tensor:
    return D.empty_colmajor(ws.Nf, ws.Nr, torch.float32, ws.device)


def _split_args(args):
    """Parse zdtfdr!([w,] [dest(s),] workspace[, spectrogram])."""
    idx = [i for i, a in enumerate(args) if isinstance(a, ZDTWorkspace)]
    if len(idx) != 1:
        raise TypeError("zdtfdr_ needs exactly one ZDTWorkspace argument")
    i = idx[0]
    ws = args[i]
    before, after = list(args[:i]), list(args[i + 1:])
    if len(after) > 1:
        raise TypeError("too many arguments after the workspace")
    spec = after[0] if after else None
    w = None
    dests = None
    if len(before) == 2:
        w, dests = before
    elif len(before) == 1:
        b = before[0]
        if isinstance(b, torch.Tensor) or (isinstance(b, (list, tuple)) and b and isinstance(b[0], torch.Tensor)):
            dests = b
        else:
            w = b
    elif len(before) > 2:
        raise TypeError("too many arguments before the workspace")
    return w, dests, ws, spec


def zdtfdr_(*args, r0=None):
    """zdtfdr!([w,] [dest(s),] workspace[, spectrogram]; r0=workspace.r0)
    (src/zdtfdr.jl:349-396).  With a dest matrix returns it; with an iterable
    of dests runs one batch per (dest, rate) in zip(dests, cycle(r0)) and
    returns the dests; with no dest leaves the result in the workspace for
    ``output_`` and returns None."""
    w, dests, ws, spec = _split_args(args)
    if spec is not None:
        input_(ws, spec)
    kind, wt = ws.window(w)
    r0v = ws.r0 if r0 is None else r0
    if dests is None:
        if isinstance(r0v, (list, tuple, np.ndarray, RateRange)):
            raise TypeError("a single batch takes a scalar r0")
        _lib.check(_lib.lib().zdt_run(ws.handle, float(float32_of(r0v)), kind, _ptr(wt), None))
        return None
    if isinstance(dests, torch.Tensor):
        _check_dest(ws, dests)
        _lib.check(_lib.lib().zdt_run(ws.handle, float(float32_of(r0v)), kind, _ptr(wt), dests.data_ptr()))
        return dests
    dests = list(dests)
    rates = r0v if isinstance(r0v, (list, tuple, np.ndarray, RateRange)) else [r0v]
    r0s = []
    for d, r in zip(dests, itertools.cycle(list(rates))):
        _check_dest(ws, d)
        r0s.append(float(float32_of(r)))
    nb = len(dests)
    arr = (C.c_float * nb)(*r0s)
    ptrs = (C.c_void_p * nb)(*[d.data_ptr() for d in dests])
    _lib.check(_lib.lib().zdt_run_batches(ws.handle, arr, nb, kind, _ptr(wt), ptrs))
    return dests


def zdtfdr(*args, r0=None) -> torch.Tensor:
    """zdtfdr([w,] workspace[, spectrogram]; r0): new (Nf, Nr) FDR matrix
    (src/zdtfdr.jl:412-421)."""
    w, dests, ws, spec = _split_args(args)
    if dests is not None:
        raise TypeError("zdtfdr takes no dest; use zdtfdr_")
    dest = create_fdr_for(ws)
    if w is None:
        return zdtfdr_(dest, ws, *([spec] if spec is not None else []), r0=r0)
    return zdtfdr_(w, dest, ws, *([spec] if spec is not None else []), r0=r0)


def fdrsynchronize(ws: Optional[ZDTWorkspace] = None) -> None:
    """fdrsynchronize: wait for the device (ext/CUDAFrequencyDriftRateTransformsExt.jl:26-28)."""
    if ws is not None:
        _lib.check(_lib.lib().zdt_sync(ws.handle))
    else:
        torch.cuda.synchronize()
