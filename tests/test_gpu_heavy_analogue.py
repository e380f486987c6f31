"""The checks of the reference's network-only integration test
(test/heavytests.jl:17-40: 150 channels of a Voyager-1 spectrogram, rates
0:-0.25:-5; intfdr, fftfdr and zdtfdr peak at the same cell, and fdshift at
the carrier's drift aligns every time step's peak within two channels) on a
synthetic spectrogram of the same shape: Exponential(1) noise plus a carrier
starting between channels 54 and 55 (0-based) and drifting at -2.43
channels per step.  The HDF5 file cannot be fetched here; the peak cell is
the one the FP64 oracle and the integer-shift intfdr agree on (55, 10)
(0-based; test_heavy_analogue_oracle pins it on the CPU)."""
import sys

import numpy as np
import pytest

from oracle import synth
from oracle import zdt_oracle as Z

NF, NT = 150, 16
RATES = [-0.25 * j for j in range(21)]  # 0:-0.25:-5
PEAK = (55, 10)


def spectrogram():
    return synth.spectrogram(NF, NT, [(54.5, -2.43, 40.0)], seed=2020)


def argmax2(a):
    return tuple(int(i) for i in np.unravel_index(np.argmax(a), a.shape))


def test_heavy_analogue_oracle():
    s = spectrogram()
    assert argmax2(Z.intfdr(s, RATES)) == PEAK
    assert argmax2(Z.fftfdr64(s, RATES)) == PEAK
    assert argmax2(Z.fdr64(s, 0.0, -0.25, 21)) == PEAK
    assert set(np.argmax(Z.fdshift64(s, -2.43), axis=0).tolist()) <= {54, 55}


@pytest.mark.gpu
def test_heavy_analogue_gpu(gpu):
    import fdrt_hip as F
    print("heavy-test analogue: 150 x 16, rates 0:-0.25:-5", flush=True)
    sys.stdout.flush()
    s = spectrogram()
    ws = F.fftfdr_workspace(s)
    assert argmax2(F.fftfdr(ws, RATES).cpu().numpy()) == PEAK
    dedop = F.fdshift(ws, -2.43).cpu().numpy()
    assert set(np.argmax(dedop, axis=0).tolist()) <= {54, 55}
    zws = F.ZDTWorkspace(s, RATES)
    assert argmax2(F.zdtfdr(zws).cpu().numpy()) == PEAK
