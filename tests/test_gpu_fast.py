"""GPU parity of the fused path (power-of-two Nf in [2^16, 2^20], Nt <= 16)
against the FP64 oracle and against the generic kernels.  Same tolerances
as test_gpu_parity.py (normwise <= 2e-6 vs FP64)."""
import sys
from fractions import Fraction

import numpy as np
import pytest

from oracle import host_oracle as H
from oracle import synth
from oracle import zdt_oracle as Z

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def normwise(a, b):
    a = np.asarray(a)
    a = a.astype(np.complex128) if np.iscomplexobj(a) else a.astype(np.float64)
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def host(t):
    return t.detach().cpu().numpy()


@pytest.fixture(scope="module")
def F(gpu):
    import fdrt_hip
    return fdrt_hip


@pytest.mark.parametrize("Nf", [1 << 16, 1 << 17, 1 << 18, 1 << 19, 1 << 20])
@pytest.mark.parametrize("Nt", [16, 7, 3, 1])
def test_fast_sizes(F, Nf, Nt):
    """Every k2 time-step template (Nt padded to 16, 8, 4, 2) at every N1;
    N1 = 512 (2^20 channels) runs the q64 row layout."""
    Nr = 33 if Nf >= 1 << 19 else 65
    s = synth.spectrogram(Nf, Nt, [(Nf * 0.31, 1.5, 25.0), (Nf * 0.77, -2.75, 40.0)], seed=Nf % 1000 + Nt)
    say(f"fast Nf={Nf} Nt={Nt} Nr={Nr}")
    ws = F.ZDTWorkspace(s, -3.0, 0.1875, Nr)
    assert ws.path == "fused"
    out = host(F.zdtfdr(ws))
    ref = Z.fdr64(s, -3.0, 0.1875, Nr)
    assert normwise(out, ref) <= 2e-6


@pytest.mark.parametrize("Nf,r0,dr,Nr,Nt", [(1 << 17, -3.75, 0.5, 16, 16), (1 << 17, -4.0, 0.25, 33, 16),
                                           (1 << 17, 2.0, -0.25, 17, 11), (1 << 17, -0.5, 1.0, 2, 16),
                                           (1 << 17, -1.0, 1.0, 3, 13), (1 << 17, -1.0, 0.5, 5, 16),
                                           (1 << 19, -2.0, 0.125, 33, 16), (1 << 19, -1.25, 0.25, 11, 14),
                                           (1 << 17, 0.0, -0.25, 21, 16), (1 << 17, 0.5, 0.25, 20, 12),
                                           (1 << 18, -4.0, 1.0 / 60, 481, 16), (1 << 19, 1.0, 0.0625, 64, 16)])
def test_fast_paired_rates(F, Nf, r0, dr, Nr, Nt):
    """k2 pairs rates j and Nr-1-j around the centre rate: even Nr, odd Nr (a
    middle rate), descending rates, one pair, one pair and the middle, a last
    group holding only the middle rate, 16-row groups (4 pairs) at 2^19, and
    rate sets not symmetric about 0 (all negative, all positive, a
    non-dyadic step)."""
    s = synth.spectrogram(Nf, Nt, [(Nf * 0.23, 0.75, 30.0), (Nf * 0.66, -1.5, 25.0)], seed=Nr + Nt)
    say(f"fast paired r0={r0} dr={dr} Nr={Nr} Nt={Nt}")
    out = host(F.zdtfdr(F.ZDTWorkspace(s, r0, dr, Nr)))
    assert normwise(out, Z.fdr64(s, r0, dr, Nr)) <= 2e-6


@pytest.mark.parametrize("Nf", [1 << 17, 1 << 18, 1 << 19])
def test_fast_n2_512(F, Nf):
    Nr = 33
    s = synth.spectrogram(Nf, 16, [(Nf * 0.43, 2.5, 25.0), (Nf * 0.12, -0.75, 35.0)], seed=Nf % 997)
    say(f"fast N2=512 Nf={Nf}")
    out = host(F.zdtfdr(F.ZDTWorkspace(s, -3.0, 0.1875, Nr, n2=512)))
    assert normwise(out, Z.fdr64(s, -3.0, 0.1875, Nr)) <= 2e-6


def test_fast_n2_512_unavailable(F):
    s = synth.spectrogram(1 << 20, 4, [(1000.0, 1.0, 20.0)], seed=3)
    with pytest.raises(F.ZDTError):
        F.ZDTWorkspace(s, -1.0, 0.5, 5, n2=512)


@pytest.mark.parametrize("Nt", [2, 4, 8, 12, 16])
def test_fast_vs_generic(F, Nt):
    Nf = 1 << 17
    s = synth.spectrogram(Nf, Nt, [(5000.0, 0.5, 30.0)], seed=40 + Nt)
    say(f"fast vs generic Nt={Nt}")
    wa = F.ZDTWorkspace(s, -4.0, 0.0625, 129)
    wb = F.ZDTWorkspace(s, -4.0, 0.0625, 129, generic=True)
    assert (wa.path, wb.path) == ("fused", "generic")
    a = host(F.zdtfdr(wa))
    b = host(F.zdtfdr(wb))
    assert normwise(a, b.astype(np.float64)) <= 2e-6


@pytest.mark.parametrize("w", ["hamming", "custom"])
def test_fast_windows(F, w):
    Nf = 1 << 16
    s = synth.spectrogram(Nf, 16, [(1000.0, 0.75, 20.0)], seed=4)
    fn = (lambda n, N: 0.5 + 0.5 * np.cos(np.pi * n / N))
    win = "hamming" if w == "hamming" else fn
    say(f"fast window {w}")
    ws = F.ZDTWorkspace(s, -2.0, 0.125, 33)
    out = host(F.zdtfdr(win, ws))
    ref = Z.fdr64(s, -2.0, 0.125, 33, window=win)
    assert normwise(out, ref) <= 2e-6


def test_fast_batches_output_and_F(F):
    Nf = 1 << 18
    s = synth.spectrogram(Nf, 16, [(70000.0, 1.5, 30.0), (200000.0, -3.25, 25.0)], seed=8)
    say("fast batches / no-dest / F")
    ws = F.ZDTWorkspace(s, -4.0, 0.0625, 43)
    r0s = [-4.0, -4.0 + 43 * 0.0625, -4.0 + 86 * 0.0625]
    dests = [F.create_fdr(s, 43) for _ in r0s]
    F.zdtfdr_(dests, ws, r0=r0s)
    full = np.concatenate([host(d) for d in dests], axis=1)
    assert normwise(full, Z.fdr64(s, -4.0, 0.0625, 129)) <= 2e-6
    assert F.zdtfdr_(ws, r0=-1.0) is None
    dest = F.create_fdr(s, 43)
    F.output_(dest, ws)
    F.output_(dest, ws)  # output! may be repeated (the spectra are not clobbered)
    assert normwise(host(dest), Z.fdr64(s, -1.0, 0.0625, 43)) <= 2e-6
    Fk = host(ws.F)
    ref = np.fft.rfft(s.astype(np.float64), axis=0)
    assert normwise(Fk, ref) <= 1e-6
    # step functions on a fused workspace
    F.preprocess_(ws, -4.0)
    F.convolve_(ws)
    F.postprocess_(ws)
    F.output_(dest, ws)
    assert normwise(host(dest), Z.fdr64(s, -4.0, 0.0625, 43)) <= 2e-6


def test_fast_unnormalized_and_hits(F):
    Nf = 1 << 16
    s = synth.cfg1(seed=77)
    say("fast unnormalized + hits")
    ws = F.ZDTWorkspace(s, -4.0, 0.25, 33, unnormalized=True)
    out = host(F.zdtfdr(ws)) / (Nf * ws.Nl)
    ref = Z.fdr64(s, -4.0, 0.25, 33)
    assert normwise(out, ref) <= 2e-6
    ws2 = F.ZDTWorkspace(s, -4.0, 0.25, 33)
    fdr = F.zdtfdr(ws2)
    np.testing.assert_array_equal(F.findprotohits(fdr, 10, snr=True), H.findprotohits(host(fdr), 10, snr=True))


# ---- per-bin chirp-Z with the fused output (power-of-two Nf, Nt > 16)

@pytest.mark.parametrize("Nf,Nt", [(1 << 16, 17), (1 << 17, 32), (1 << 18, 23), (1 << 19, 60), (1 << 20, 33),
                                   (1 << 16, 64), (1 << 16, 150), (1 << 17, 100)])
def test_fused_output_vs_fp64(F, Nf, Nt):
    # rates -2:1/8:2 pair as r_j = -r_{32-j} (the paired-rate evaluation)
    Nr = 33
    s = synth.spectrogram(Nf, Nt, [(Nf * 0.37, 1.25, 25.0), (Nf * 0.81, -2.0, 30.0)], seed=Nf % 991 + Nt)
    say(f"fused output Nf={Nf} Nt={Nt}")
    ws = F.ZDTWorkspace(s, -2.0, 0.125, Nr)
    assert ws.path == "fused-output"
    out = host(F.zdtfdr(ws))
    assert normwise(out, Z.fdr64(s, -2.0, 0.125, Nr)) <= 2e-6


@pytest.mark.parametrize("Nf,Nt,r0,dr,Nr", [(1 << 16, 1024, -0.5, 1.0 / 64, 5), (1 << 18, 128, -1.0, 1.0 / 32, 9),
                                            (1 << 22, 80, -3.0, 1.5, 3), (1 << 21, 65, 0.25, 0.5, 4)])
def test_fused_output_long_integrations(F, Nf, Nt, r0, dr, Nr):
    """The chunked column chirp-Z (64-step chunks) at its Nt = 1024 limit,
    at 2^18 x 128, and at N1 = 2048 and 1024 (2^22 and 2^21 channels) with
    more than 64 steps."""
    s = synth.spectrogram(Nf, Nt, [(Nf * 0.29, 0.375, 8.0), (Nf * 0.71, -1.0, 6.0)], seed=Nt + Nr)
    say(f"fused output long Nf={Nf} Nt={Nt} Nr={Nr}")
    ws = F.ZDTWorkspace(s, r0, dr, Nr)
    assert ws.path == "fused-output"
    out = host(F.zdtfdr(ws))
    assert normwise(out, Z.fdr64(s, r0, dr, Nr)) <= 2e-6


@pytest.mark.parametrize("Nf,Nt", [(6000, 1100), (1 << 16, 1100)])
def test_generic_many_steps(F, Nf, Nt):
    """More than 1024 steps (past the fused output's limit) and a
    non-power-of-two Nf: the generic path's chunked chirp-Z, 18 chunks."""
    r0, dr, Nr = -0.25, 1.0 / 128, 7
    s = synth.spectrogram(Nf, Nt, [(Nf * 0.4, 0.125, 5.0)], seed=Nf % 97)
    say(f"generic many steps Nf={Nf} Nt={Nt}")
    ws = F.ZDTWorkspace(s, r0, dr, Nr)
    assert ws.path == "generic"
    assert normwise(host(F.zdtfdr(ws)), Z.fdr64(s, r0, dr, Nr)) <= 2e-6


def test_fused_output_chunked_asymmetric_rates(F):
    """Nt = 129 (chirp-Z chunks of 64, 64 and 1 steps) with rates not
    symmetric about 0 (centre rate 4), against the oracle and the generic
    kernels."""
    Nf, Nt, r0, dr, Nr = 1 << 16, 129, 1.5, 0.25, 21
    s = synth.spectrogram(Nf, Nt, [(Nf * 0.3, 3.0, 25.0), (Nf * 0.6, 5.25, 20.0)], seed=129)
    say("fused output Nt=129, rates 1.5:0.25:6.5")
    out = host(F.zdtfdr(F.ZDTWorkspace(s, r0, dr, Nr)))
    assert normwise(out, Z.fdr64(s, r0, dr, Nr)) <= 2e-6
    g = host(F.zdtfdr(F.ZDTWorkspace(s, r0, dr, Nr, generic=True)))
    assert normwise(out, g.astype(np.float64)) <= 2e-6


def test_fused_output_window_nodest_unnormalized(F):
    Nf, Nt = 1 << 17, 40
    s = synth.spectrogram(Nf, Nt, [(40000.0, 0.75, 30.0)], seed=17)
    say("fused output: hamming, no-dest + output!, unnormalized")
    ws = F.ZDTWorkspace(s, -1.0, 0.0625, 33)
    ref = Z.fdr64(s, -1.0, 0.0625, 33, window="hamming")
    assert normwise(host(F.zdtfdr("hamming", ws)), ref) <= 2e-6
    assert F.zdtfdr_(ws, r0=-1.0) is None
    dest = F.create_fdr(s, 33)
    F.output_(dest, ws)
    F.output_(dest, ws)
    assert normwise(host(dest), Z.fdr64(s, -1.0, 0.0625, 33)) <= 2e-6
    wu = F.ZDTWorkspace(s, -1.0, 0.0625, 33, unnormalized=True)
    assert normwise(host(F.zdtfdr(wu)) / (Nf * wu.Nl), Z.fdr64(s, -1.0, 0.0625, 33)) <= 2e-6
    g = host(F.zdtfdr(F.ZDTWorkspace(s, -1.0, 0.0625, 33, generic=True)))
    assert normwise(host(F.zdtfdr(ws)), g.astype(np.float64)) <= 2e-6


@pytest.mark.parametrize("r0,dr,Nr", [(-2.0, 0.125, 45), (-1.0, 0.125, 40), (0.5, 0.25, 20), (3.0, -0.25, 25),
                                     (-0.375, 0.25, 4), (-0.3, 0.25, 20), (1.5, 0.25, 1), (-4.0, 1.0 / 60, 61)])
def test_fused_output_rate_sets(F, r0, dr, Nr):
    """The paired evaluation around the centre rate on rate sets that are
    not symmetric about 0 (shifted, all positive, descending, non-dyadic
    steps), on a symmetric one of 4 rates and on a single rate."""
    Nf, Nt = 1 << 17, 24
    s = synth.spectrogram(Nf, Nt, [(Nf * 0.41, 1.75, 25.0), (Nf * 0.77, -0.5, 30.0)], seed=Nr)
    say(f"fused output rates r0={r0} dr={dr} Nr={Nr}")
    out = host(F.zdtfdr(F.ZDTWorkspace(s, r0, dr, Nr)))
    assert normwise(out, Z.fdr64(s, r0, dr, Nr)) <= 2e-6


@pytest.mark.parametrize("Nf,Nt", [(1 << 21, 17), (1 << 22, 20)])
def test_fused_output_large_nf(F, Nf, Nt):
    s = synth.spectrogram(Nf, Nt, [(Nf * 0.62, 3.0, 30.0), (Nf * 0.05, -1.5, 25.0)], seed=Nt)
    say(f"fused output Nf={Nf} Nt={Nt}")
    out = host(F.zdtfdr(F.ZDTWorkspace(s, -1.5, 1.5, 3)))
    assert normwise(out, Z.fdr64(s, -1.5, 1.5, 3)) <= 2e-6


def test_fused_output_F_reinput_and_step_functions(F):
    Nf, Nt = 1 << 18, 24
    s1 = synth.spectrogram(Nf, Nt, [(90000.0, 1.25, 30.0)], seed=23)
    s2 = synth.spectrogram(Nf, Nt, [(150000.0, -2.5, 25.0)], seed=24)
    say("fused output: F buffer, input! again, step functions")
    ws = F.ZDTWorkspace(s1, -2.0, 0.125, 33)
    ref1 = np.fft.rfft(s1.astype(np.float64), axis=0)
    assert normwise(host(ws.F), ref1) <= 1e-6
    # a new input! refreshes the unpacked F and the transform
    F.input_(ws, s2)
    assert normwise(host(ws.F), np.fft.rfft(s2.astype(np.float64), axis=0)) <= 1e-6
    assert normwise(host(F.zdtfdr(ws)), Z.fdr64(s2, -2.0, 0.125, 33)) <= 2e-6
    # the step functions read the natural-layout F
    dest = F.create_fdr(s2, 33)
    F.preprocess_(ws, -2.0)
    F.convolve_(ws)
    F.postprocess_(ws)
    F.output_(dest, ws)
    assert normwise(host(dest), Z.fdr64(s2, -2.0, 0.125, 33)) <= 2e-6
    # and the fused output again after them
    assert normwise(host(F.zdtfdr(ws, s1)), Z.fdr64(s1, -2.0, 0.125, 33)) <= 2e-6
