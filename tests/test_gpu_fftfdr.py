"""GPU parity of the Fourier-shift de-doppler entry points (zdt_fftfdr,
zdt_fdshift; src/fftfdr.jl) against the FP64 oracle."""
import ctypes as C
import sys

import numpy as np
import pytest

from oracle import synth
from oracle import zdt_oracle as Z

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def normwise(a, b):
    return float(np.max(np.abs(a.astype(np.float64) - b)) / np.max(np.abs(b)))


def host(t):
    return t.detach().cpu().numpy()


@pytest.fixture(scope="module")
def F(gpu):
    import fdrt_hip
    return fdrt_hip


RATES = [-3.7, -1.0, 0.0, 0.33, 1.5, 2.0, 3.99, 7.25]


@pytest.mark.parametrize("Nf,Nt", [(4096, 16), (6000, 40), (1 << 16, 7), (2048, 70)])
def test_fftfdr_vs_fp64(F, Nf, Nt):
    s = synth.spectrogram(Nf, Nt, [(Nf * 0.4, 1.5, 25.0), (Nf * 0.8, -3.7, 30.0)], seed=Nf + Nt)
    say(f"fftfdr Nf={Nf} Nt={Nt}")
    ws = F.fftfdr_workspace(s)
    out = host(F.fftfdr(ws, RATES))
    assert normwise(out, Z.fftfdr64(s, RATES)) <= 2e-6
    col = host(F.fdshiftsum(ws, 1.5))
    assert normwise(col, Z.fftfdr64(s, [1.5])[:, 0]) <= 2e-6


@pytest.mark.parametrize("Nf,Nt", [(4096, 16), (6000, 9)])
def test_fdshift_vs_fp64(F, Nf, Nt):
    s = synth.spectrogram(Nf, Nt, [(Nf * 0.25, -2.25, 25.0)], seed=7 + Nt)
    say(f"fdshift Nf={Nf} Nt={Nt}")
    ws = F.fftfdr_workspace(s)
    for rate in (-2.25, 0.5):
        out = host(F.fdshift(ws, rate))
        assert out.shape == (Nf, Nt)
        assert normwise(out, Z.fdshift64(s, rate)) <= 2e-6
    # the de-dopplered tone lines up in one channel: summing over time is fdshiftsum
    sh = host(F.fdshift(ws, -2.25)).astype(np.float64).sum(axis=1)
    assert normwise(sh, host(F.fdshiftsum(ws, -2.25)).astype(np.float64)) <= 2e-6


def test_fftfdr_equals_zdt_on_a_range(F):
    Nf = 1 << 16
    s = synth.cfg1(seed=12)
    ws = F.fftfdr_workspace(s)
    rates = [-4.0 + 0.25 * j for j in range(33)]
    a = host(F.fftfdr(ws, rates))
    b = host(F.zdtfdr(F.ZDTWorkspace(s, -4.0, 0.25, 33)))
    assert normwise(a, b.astype(np.float64)) <= 2e-6
    # a new spectrogram through fftfdr_workspace! replaces the spectrum
    s2 = synth.spectrogram(Nf, 16, [(30000.0, 2.0, 40.0)], seed=5)
    F.fftfdr_workspace_(ws, s2)
    assert normwise(host(F.fftfdr(ws, [2.0])), Z.fftfdr64(s2, [2.0])) <= 2e-6


def test_fftfdr_on_fused_workspace(F):
    import torch
    Nf = 1 << 17
    s = synth.spectrogram(Nf, 16, [(Nf * 0.3, 0.75, 30.0)], seed=21)
    ws = F.ZDTWorkspace(s, -1.0, 0.25, 9)  # fused path: F is unpacked on demand
    r = np.asarray(RATES, dtype=np.float32)
    dest = torch.empty((len(r), Nf), device="cuda").t()
    from fdrt_hip import _lib
    _lib.check(_lib.lib().zdt_fftfdr(ws.handle, r.ctypes.data_as(C.POINTER(C.c_float)), len(r), dest.data_ptr()))
    assert normwise(host(dest), Z.fftfdr64(s, RATES)) <= 2e-6


def test_fftfdr_errors(F):
    s = synth.spectrogram(1024, 4, [(100.0, 1.0, 20.0)], seed=1)
    ws = F.fftfdr_workspace(s)
    import torch
    with pytest.raises(ValueError):
        F.fftfdr_(torch.empty((3, 1024), device="cuda").t(), ws, [1.0])
    with pytest.raises(F.ZDTError):
        F.fdshift(ws, float("nan"))
