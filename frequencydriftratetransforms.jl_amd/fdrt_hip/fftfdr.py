"""Fourier-shift de-doppler (src/fftfdr.jl) over libzdt_hip.so: fdshift,
fdshiftsum and fftfdr for any list of rates.

``fftfdr_workspace(spectrogram)`` holds the input spectrum (the reference's
``dest_rfft``, src/fftfdr.jl:39-59); the transforms read it on the device.
Phases are reduced exactly (the reference rounds ``2*k*n*r`` to Float32 in
``phasor``, src/fftfdr.jl:10-12).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _device as D
from . import _lib
from .zdt import ZDTWorkspace, input_


class FftFdrWorkspace:
    """fftfdr_workspace(spectrogram) -- src/fftfdr.jl:39-59."""

    def __init__(self, spectrogram):
        # one rate, generic kernels: the workspace only serves its input spectrum
        self.ws = ZDTWorkspace(spectrogram, 0.0, 1.0, 1, generic=True)
        self.Nf, self.Nt = self.ws.Nf, self.ws.Nt

    @property
    def dest_rfft(self) -> torch.Tensor:
        return self.ws.F


def fftfdr_workspace(spectrogram, bunaligned: bool = True) -> FftFdrWorkspace:
    return FftFdrWorkspace(spectrogram)


def fftfdr_workspace_(workspace, spectrogram, bunaligned: bool = True) -> FftFdrWorkspace:
    """fftfdr_workspace!(workspace, spectrogram) -- src/fftfdr.jl:69-84."""
    if workspace is None:
        return FftFdrWorkspace(spectrogram)
    spec = D.colmajor(spectrogram)
    if tuple(spec.shape) != (workspace.Nf, workspace.Nt):
        raise ValueError("input spectrogram has unexpected size")
    input_(workspace.ws, spec)
    return workspace


def fftfdr_(fdr: torch.Tensor, workspace: FftFdrWorkspace, rates) -> torch.Tensor:
    """fftfdr!(fdr, workspace, rates) -- src/fftfdr.jl:126-134."""
    r = np.asarray(list(rates), dtype=np.float32)
    if tuple(fdr.shape) != (workspace.Nf, len(r)) or not D.is_colmajor(fdr) or fdr.dtype != torch.float32:
        raise ValueError(f"fdr must be a column-major float32 ({workspace.Nf}, {len(r)}) tensor")
    _lib.check(_lib.lib().zdt_fftfdr(workspace.ws.handle, r.ctypes.data_as(C.POINTER(C.c_float)), len(r),
                                     fdr.data_ptr()))
    return fdr


def fftfdr(workspace: FftFdrWorkspace, rates) -> torch.Tensor:
    """fftfdr(workspace, rates) -- src/fftfdr.jl:143-148."""
    rates = list(rates)
    return fftfdr_(D.empty_colmajor(workspace.Nf, len(rates)), workspace, rates)


def fdshiftsum_(dest: torch.Tensor, workspace: FftFdrWorkspace, rate) -> torch.Tensor:
    """fdshiftsum!(dest, workspace, rate) -- src/fftfdr.jl:86-96."""
    col = dest.reshape(-1, 1) if dest.dim() == 1 else dest
    fftfdr_(col, workspace, [rate])
    return dest


def fdshiftsum(workspace: FftFdrWorkspace, rate) -> torch.Tensor:
    """fdshiftsum(workspace, rate) -- src/fftfdr.jl:98-102."""
    return fftfdr(workspace, [rate])[:, 0]


def fdshift_(dest: torch.Tensor, workspace: FftFdrWorkspace, rate) -> torch.Tensor:
    """fdshift!(dest, workspace, rate) -- src/fftfdr.jl:104-111."""
    if tuple(dest.shape) != (workspace.Nf, workspace.Nt) or not D.is_colmajor(dest) or dest.dtype != torch.float32:
        raise ValueError(f"dest must be a column-major float32 ({workspace.Nf}, {workspace.Nt}) tensor")
    _lib.check(_lib.lib().zdt_fdshift(workspace.ws.handle, float(np.float32(rate)), dest.data_ptr()))
    return dest


def fdshift(workspace: FftFdrWorkspace, rate) -> torch.Tensor:
    """fdshift(workspace, rate) -- src/fftfdr.jl:113-118."""
    return fdshift_(D.empty_colmajor(workspace.Nf, workspace.Nt), workspace, rate)
