"""GPU parity of the ZDT path through libzdt_hip.so against the oracle.

Tolerances (normwise = max|gpu - ref| / max|ref|):
  * vs the FP64 definition oracle: <= 2e-6 (FP32 FFTs of up to 2^20 points,
    exact phases);
  * vs the FP32 restatement of the reference: <= 3e-5 (dominated by the
    reference's own Float32 phase-argument rounding, src/zdtfdr.jl:141,197,269);
  * integer rates vs the integer-shift closed form (intfdr): <= 2e-6.
Proto-hit sets must be identical to the FP64 oracle's except for cells whose
FP64 value lies within eps = max(4 ulp(thr), max|gpu - fp64|) of the threshold.
"""
import json
import sys
from fractions import Fraction
from pathlib import Path

import numpy as np
import pytest

from oracle import host_oracle as H
from oracle import synth
from oracle import zdt_oracle as Z

pytestmark = pytest.mark.gpu

GOLD = json.loads((Path(__file__).parent / "golden" / "runtests_3x2.json").read_text())


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


# ------------------------------------------------------------ known answer

@pytest.mark.parametrize("generic", [False, True])
def test_runtests_known_answer(F, generic):
    d = np.asarray(GOLD["d"], np.float32)
    say(f"runtests 3x2 generic={generic}")
    ws = F.ZDTWorkspace(d, range(-1, 2), generic=generic)
    assert ws.Nl == GOLD["Nl"]
    out = host(F.zdtfdr(ws))
    np.testing.assert_allclose(out, GOLD["fdr_expected"], atol=1e-6)
    # the test's own unnormalized convention: zdtfdr(ws) ./ 12 == fdr_expected
    wsu = F.ZDTWorkspace(d, range(-1, 2), unnormalized=True, generic=generic)
    np.testing.assert_allclose(host(F.zdtfdr(wsu)) / GOLD["zdtfdr_scale_in_test"], GOLD["fdr_expected"], atol=1e-6)


def test_runtests_stepwise(F):
    d = np.asarray(GOLD["d"], np.float32)
    say("runtests 3x2 stepwise")
    ws = F.ZDTWorkspace(d, range(-1, 2))
    F.preprocess_(ws)
    F.convolve_(ws)
    F.postprocess_(ws)
    dest = F.create_fdr(d, 3)
    F.output_(dest, ws)
    np.testing.assert_allclose(host(dest), GOLD["fdr_expected"], atol=1e-6)


# ------------------------------------------------------------ vs FP64 oracle

CASES = [
    # Nf, Nt, r0, dr, Nr
    (4096, 16, -4.0, 0.0625, 129),
    (4095, 16, -4.0, 0.0625, 129),
    (1000, 7, -2.0, 0.3, 20),
    (65536, 16, -4.0, 0.25, 33),
    (64, 1, 0.5, 0.5, 3),
    (256, 40, -2.0, 0.1, 50),
    (2048, 60, -4.0, 1.0 / 60.0, 481),
    (12, 3, -1.0, 0.5, 5),
    (1 << 14, 32, -4.0, 1.0 / 32, 257),
    (6, 2, 0.0, 1.0, 1),
    (8192, 64, -2.0, 1.0 / 16, 65),
    (4096, 70, -1.0, 0.125, 17),
    (6000, 150, 0.5, 0.25, 21),  # three 64-step chunks of the chirp-Z
    # four-step frequency FFTs: M = 96 x 512, 100 x 500 and 81 x 243 (partial tiles)
    (98304, 16, -4.0, 0.0625, 129),
    (100000, 16, -2.0, 0.125, 33),
    (39366, 8, -1.0, 0.25, 9),
]


@pytest.mark.parametrize("Nf,Nt,r0,dr,Nr", CASES)
@pytest.mark.parametrize("generic", [False, True])
def test_zdt_vs_fp64(F, Nf, Nt, r0, dr, Nr, generic):
    s = synth.spectrogram(Nf, Nt, [(Nf * 0.3, 1.25, 20.0), (Nf * 0.7, -2.5, 15.0)], seed=Nf + Nt)
    say(f"zdt Nf={Nf} Nt={Nt} r0={r0} dr={dr} Nr={Nr} generic={generic}")
    ws = F.ZDTWorkspace(s, r0, dr, Nr, generic=generic)
    out = host(F.zdtfdr(ws))
    ref = Z.fdr64(s, r0, dr, Nr)
    assert out.shape == (Nf, Nr)
    assert normwise(out, ref) <= 2e-6


@pytest.mark.parametrize("Nf,Nt,r0,dr,Nr", [(98304, 16, -4.0, 0.0625, 129), (39366, 8, -1.0, 0.25, 9),
                                             (1 << 15, 12, -2.0, 0.25, 17)])
def test_zdt_generic_radix8_vs_fp64(F, Nf, Nt, r0, dr, Nr):
    """The generic FFTs with radices up to 8 (a ragged plan's other candidate)."""
    s = synth.spectrogram(Nf, Nt, [(Nf * 0.4, 0.75, 20.0)], seed=Nf + 3)
    say(f"zdt radix 8 Nf={Nf} Nt={Nt}")
    ws = F.ZDTWorkspace(s, r0, dr, Nr, generic=True, radix_cap=8)
    assert normwise(host(F.zdtfdr(ws)), Z.fdr64(s, r0, dr, Nr)) <= 2e-6


@pytest.mark.parametrize("Nf,Nt,r0,dr,Nr", [(3 << 19, 4, -1.0, 0.5, 5), (3 << 20, 2, 0.0, 1.0, 2)])
def test_zdt_generic_fft_limits(F, Nf, Nt, r0, dr, Nr):
    """The largest four-step frequency FFTs (M = 768 x 1024: 8-column tiles,
    a 768-entry high twiddle table) and the pass-per-radix FFTs above them
    (M = 3 x 2^19 > 2^20)."""
    s = synth.spectrogram(Nf, Nt, [(Nf * 0.25, 0.5, 30.0)], seed=Nf % 977)
    say(f"zdt generic FFT limits Nf={Nf} Nt={Nt}")
    ws = F.ZDTWorkspace(s, r0, dr, Nr, generic=True)
    assert normwise(host(F.zdtfdr(ws)), Z.fdr64(s, r0, dr, Nr)) <= 2e-6


@pytest.mark.parametrize("generic", [False, True])
def test_integer_rates_vs_intfdr(F, generic):
    s = synth.cfg1(seed=21)
    say(f"intfdr cfg1 integer rates generic={generic}")
    ws = F.ZDTWorkspace(s, range(-4, 5), generic=generic)
    out = host(F.zdtfdr(ws))
    ref = Z.intfdr(s.astype(np.float64), range(-4, 5))
    assert normwise(out, ref) <= 2e-6


def test_vs_reference_fp32_restatement(F):
    s = synth.cfg1(seed=5)
    say("cfg1 vs ref32")
    ws = F.ZDTWorkspace(s, -4.0, 0.25, 33)
    out = host(F.zdtfdr(ws))
    ref32 = Z.zdt_ref32(s, -4.0, 0.25, 33)
    assert normwise(out, ref32.astype(np.float64)) <= 3e-5


@pytest.mark.parametrize("w", ["hamming", "custom", "vector"])
def test_windows(F, w):
    Nf = 8192
    s = synth.spectrogram(Nf, 16, [(1000.0, 0.75, 20.0)], seed=4)
    fn = (lambda n, N: 0.5 + 0.5 * np.cos(np.pi * n / N))
    if w == "hamming":
        win, ref = "hamming", Z.fdr64(s, -2.0, 0.125, 33, window="hamming")
    elif w == "custom":
        win, ref = fn, Z.fdr64(s, -2.0, 0.125, 33, window=fn)
    else:
        vals = np.array([fn(n, Nf) for n in range(Nf // 2 + 1)])
        win, ref = vals.astype(np.float32), Z.fdr64(s, -2.0, 0.125, 33, window=vals.astype(np.float32).astype(np.float64))
    say(f"window {w}")
    ws = F.ZDTWorkspace(s, -2.0, 0.125, 33)
    out = host(F.zdtfdr(win, ws))
    assert normwise(out, ref) <= 2e-6
    # the same window through the step functions
    F.preprocess_(ws)
    F.convolve_(ws)
    F.postprocess_(win, ws)
    dest = F.create_fdr(s, 33)
    F.output_(dest, ws)
    assert normwise(host(dest), ref) <= 2e-6


def test_unsupported_window(F):
    s = synth.spectrogram(64, 4, [], seed=1)
    ws = F.ZDTWorkspace(s, 0.0, 1.0, 3)
    with pytest.raises(ValueError, match="unsupported window type"):
        F.zdtfdr("blackman", ws)
    with pytest.raises(ValueError, match="entries"):
        F.zdtfdr(np.ones(10, np.float32), ws)


def test_r0_override_and_batches(F):
    Nf = 16384
    s = synth.spectrogram(Nf, 16, [(4000.0, 1.5, 30.0), (9000.0, -3.25, 25.0)], seed=8)
    say("r0 override / multi-batch")
    ws = F.ZDTWorkspace(s, -4.0, 0.0625, 43)
    out = host(F.zdtfdr(ws, r0=-1.3125))
    assert normwise(out, Z.fdr64(s, -1.3125, 0.0625, 43)) <= 2e-6
    r0s = [-4.0, -4.0 + 43 * 0.0625, -4.0 + 86 * 0.0625]
    dests = [F.create_fdr(s, 43) for _ in r0s]
    F.zdtfdr_(dests, ws, r0=r0s)
    full = np.concatenate([host(d) for d in dests], axis=1)
    ref = Z.fdr64(s, -4.0, 0.0625, 129)
    assert normwise(full, ref) <= 2e-6
    # a scalar r0 cycles over the dests (zip(dests, cycle(r0)), src/zdtfdr.jl:355)
    d2 = [F.create_fdr(s, 43) for _ in range(2)]
    F.zdtfdr_(d2, ws, r0=-1.0)
    np.testing.assert_array_equal(host(d2[0]), host(d2[1]))
    # new input through zdtfdr_(dest, ws, spectrogram)
    s2 = synth.spectrogram(Nf, 16, [(100.0, 2.0, 40.0)], seed=9)
    dest = F.create_fdr(s2, 43)
    F.zdtfdr_(dest, ws, s2)
    assert normwise(host(dest), Z.fdr64(s2, -4.0, 0.0625, 43)) <= 2e-6


def test_no_dest_then_output(F):
    s = synth.spectrogram(4096, 16, [(1000.0, 1.0, 20.0)], seed=3)
    say("no-dest run then output_")
    ws = F.ZDTWorkspace(s, -4.0, 0.125, 65)
    assert F.zdtfdr_(ws) is None
    dest = F.create_fdr(s, 65)
    F.output_(dest, ws)
    assert normwise(host(dest), Z.fdr64(s, -4.0, 0.125, 65)) <= 2e-6


def test_stepwise_matches_reference_steps(F):
    Nf, Nt = 2048, 16
    s = synth.spectrogram(Nf, Nt, [(300.0, 1.75, 20.0)], seed=12)
    say("stepwise vs ref32 steps")
    ws = F.ZDTWorkspace(s, -4.0, 0.0625, 129)
    ref = Z.RefWorkspace32(s, -4.0, 0.0625, 129)
    np.testing.assert_allclose(host(ws.F), ref.F, rtol=0, atol=2e-6 * np.abs(ref.F).max())
    F.preprocess_(ws)
    ref.preprocess()
    assert normwise(host(ws.Y), ref.Y.astype(np.complex128)) <= 3e-5
    F.convolve_(ws)
    ref.convolve()
    assert normwise(host(ws.Y), ref.Y.astype(np.complex128)) <= 3e-5
    F.postprocess_(ws)
    ref.postprocess()
    assert normwise(host(ws.Ys), ref.Ys.astype(np.complex128)) <= 3e-5
    dest = F.create_fdr(s, 129)
    F.output_(dest, ws)
    assert normwise(host(dest), Z.fdr64(s, -4.0, 0.0625, 129)) <= 2e-6
    # V is the FFT of the chirp kernel, computed in FP64
    assert normwise(host(ws.V), ref.V.astype(np.complex128)) <= 3e-5


def test_shape_errors(F):
    s = synth.spectrogram(64, 4, [], seed=1)
    ws = F.ZDTWorkspace(s, 0.0, 1.0, 3)
    with pytest.raises(ValueError, match="size"):
        F.input_(ws, np.zeros((32, 4), np.float32))
    with pytest.raises(ValueError, match="column-major"):
        F.output_(F.create_fdr(s, 4), ws)


# ------------------------------------------------------------ statistics and hits

def test_stats_normalize_hits_vs_oracle(F):
    import torch
    s = synth.cfg1(seed=31)
    say("stats/hits cfg1")
    ws = F.ZDTWorkspace(s, -4.0, 0.25, 33)
    fdr = F.zdtfdr(ws)
    a = host(fdr)
    m, sd = F.fdrstats(fdr)
    mo, so = H.fdrstats(a)
    assert abs(m - mo) <= 1e-12 * abs(mo) + 1e-12 and abs(sd - so) <= 1e-12 * so
    thr = F.fdrdenormalize(10, fdr)
    assert thr == pytest.approx(10 * so + mo, rel=1e-12)
    hits = F.findprotohits(fdr, 10, snr=True)
    ref = H.findprotohits(a, 10, snr=True)
    np.testing.assert_array_equal(hits, ref)
    assert len(hits) > 0
    # the three tones show up at their rates
    rates = np.arange(-4, 4.0001, 0.25)
    found = set(np.round(rates[hits[:, 1] - 1], 3).tolist())
    for r in (0.5, -1.25, 3.0):
        assert r in found
    # normalize in place
    F.fdrnormalize_(fdr)
    np.testing.assert_allclose(host(fdr), (a.astype(np.float64) - mo) / so, atol=2e-6)
    # no hits above the maximum
    assert F.findprotohits(fdr, float(torch.max(fdr)) + 1.0).shape == (0, 2)


def test_hits_multibatch_and_capacity(F):
    import ctypes as C

    import torch
    s = synth.cfg1(seed=32)
    say("hits multi-batch")
    ws = F.ZDTWorkspace(s, -4.0, 0.25, 11)
    dests = [F.create_fdr(s, 11) for _ in range(3)]
    F.zdtfdr_(dests, ws, r0=[-4.0, -4.0 + 11 * 0.25, -4.0 + 22 * 0.25])
    arrs = [host(d) for d in dests]
    thr = H.fdrdenormalize(8, *H.fdrstats(arrs))
    hits = F.findprotohits(dests, 8, snr=True, capacity=3)
    np.testing.assert_array_equal(hits, H.findprotohits(arrs, thr))
    m, sd = F.fdrstats(dests)
    mo, so = H.fdrstats(arrs)
    assert abs(sd - so) <= 1e-12 * so and abs(m - mo) <= 1e-12 * abs(mo) + 1e-12
    # direct C call with a small capacity returns ZDT_ETRUNC and the total
    from fdrt_hip import _lib
    buf = torch.empty(2, dtype=torch.int64, device=dests[0].device)
    arr = (C.c_void_p * 3)(*[d.data_ptr() for d in dests])
    n = C.c_int64()
    rc = _lib.lib().zdt_findprotohits(arr, 3, 65536, 11, 65536, thr, buf.data_ptr(), 2, C.byref(n), None)
    torch.cuda.synchronize()
    assert rc == _lib.ZDT_ETRUNC and n.value == len(hits)
    lin = host(buf)
    assert (lin % 65536 + 1).tolist() == hits[:2, 0].tolist()


def test_hit_set_identity_vs_fp64(F):
    s = synth.cfg1(seed=1234)
    say("hit-set identity cfg1")
    ws = F.ZDTWorkspace(s, -4.0, 0.25, 33)
    out = host(F.zdtfdr(ws))
    ref = Z.fdr64(s, -4.0, 0.25, 33)
    m, sd = H.fdrstats(ref)
    thr64 = H.fdrdenormalize(10, m, sd)
    thr = H.fdrdenormalize(10, *H.fdrstats(out))
    eps = max(4 * np.spacing(np.float32(thr64)), np.max(np.abs(out - ref)), abs(thr - thr64))
    g = set(map(tuple, H.findprotohits(out, thr).tolist()))
    r = set(map(tuple, H.findprotohits(ref, thr64).tolist()))
    band = {(i + 1, j + 1) for i, j in zip(*np.nonzero(np.abs(ref - thr64) <= eps))}
    assert (g ^ r) <= band


# ------------------------------------------------------------ full size (the headline workload)

CFG2 = json.loads((Path(__file__).parent / "golden" / "cfg2_hits_snr8.json").read_text())


@pytest.fixture(scope="module")
def cfg2(F):
    """2^20 x 16 spectrogram (16 tones, seed 1234), rates -4:1/16:4: the
    spectrogram, its tone rates and the device FDR of the fused path."""
    from oracle.synth import random_tones
    s = synth.cfg2(seed=1234)
    tones = random_tones(1 << 20, 16, (10, 50), (-4, 4), 1234)
    say("cfg2 2^20 x 16 x 129")
    ws = F.ZDTWorkspace(s, -4.0, 0.0625, 129)
    assert ws.path == "fused" and ws.precision == "fp32"
    fdr = F.zdtfdr(ws)
    return s, [r for _, r, _ in tones], fdr


def test_full_size_columns_vs_fp64(F, cfg2):
    """Normwise <= 2e-6 against fdr64 on the first, middle and last rates and
    the two rates either side of each of the 16 injected tones' rates (tones
    may share a rate column, so at least 16 columns, not 35)."""
    s, tone_rates, fdr = cfg2
    near = {min(128, max(0, int(np.floor((r + 4.0) * 16)) + d)) for r in tone_rates for d in (0, 1)}
    cols = sorted({0, 64, 128} | near)
    assert len(cols) >= 16
    out = host(fdr[:, cols])
    rates = [Fraction(-4) + Fraction(j, 16) for j in cols]
    ref = Z.fdr64(s, None, rates=rates)
    err = normwise(out, ref)
    say(f"cfg2 {len(cols)} columns: normwise {err:.2e}")
    assert err <= 2e-6
    # a checksum-of-checksums property over all 129 columns: the DC of each
    # column equals the sum of the spectrogram (every drift line visits each
    # time step once)
    colsum = host(fdr.double().sum(dim=0))
    np.testing.assert_allclose(colsum, s.astype(np.float64).sum(), rtol=1e-6)
    np.testing.assert_allclose(CFG2["colsum_fdr64"], s.astype(np.float64).sum(), rtol=1e-9)


def test_full_size_hit_set_vs_fp64(F, cfg2):
    """findprotohits(fdr, 8; snr=true) at 2^20 x 16 x 129 against the FP64
    oracle's hit set (tests/golden/cfg2_hits_snr8.json, written by
    make_cfg2_hits.py): the symmetric difference lies inside the tie band
    |fdr64 - thr64| <= eps, eps = 4 ulp(thr64) + 2e-6 max|fdr64| (the FDR
    tolerance) + |thr - thr64|."""
    s, _, fdr = cfg2
    hits = F.findprotohits(fdr, 8, snr=True)
    thr = F.fdrdenormalize(8, fdr)
    thr64 = CFG2["thr64"]
    assert abs(thr - thr64) <= 1e-5 * thr64
    eps = 4 * float(np.spacing(np.float32(thr64))) + 2e-6 * CFG2["max_abs_fdr64"] + abs(thr - thr64)
    assert eps <= CFG2["band_abs"]
    band = {(i, j) for i, j, d in CFG2["band"] if abs(d) <= eps}
    g = set(map(tuple, hits.tolist()))
    r = set(map(tuple, CFG2["hits"]))
    say(f"cfg2 hits: {len(g)} (FP64 {len(r)}), symmetric difference {len(g ^ r)}, band {len(band)}")
    assert len(r) > 0 and (g ^ r) <= band
    # the hits are listed in column-major order, as findall returns them
    lin = (hits[:, 0] - 1) + (1 << 20) * (hits[:, 1] - 1)
    assert np.all(np.diff(lin) > 0)


def test_driftrates_and_fdrsynchronize(F):
    """driftrates(ws; r0) is range(r0, step=ws.dr, length=ws.Nr)
    (src/zdtutils.jl:62-64) with the workspace's Float32 r0/dr
    (src/zdtfdr.jl:13-14); fdrsynchronize waits for the device
    (ext/CUDAFrequencyDriftRateTransformsExt.jl:26-28)."""
    s = synth.spectrogram(4096, 8, [(1000.0, 0.5, 20.0)], seed=2)
    ws = F.ZDTWorkspace(s, -4.0, 0.0625, 129)
    r = F.driftrates(ws)
    assert len(r) == 129 and r[0] == -4 and r[-1] == 4 and r[64] == 0
    assert r.step == Fraction(1, 16)
    r2 = F.driftrates(ws, -1.5)
    assert r2[0] == Fraction(-3, 2) and r2[128] == Fraction(13, 2)
    # a step that is not dyadic is held as its Float32 value
    ws3 = F.ZDTWorkspace(s, Fraction(-8, 3), Fraction(1, 15), 81)
    r3 = F.driftrates(ws3)
    assert r3.first == Fraction(float(np.float32(np.float32(-8) / np.float32(3))))
    assert r3.step == Fraction(float(np.float32(np.float32(1) / np.float32(15))))
    assert len(r3) == 81
    # fdrsynchronize: after it, a result written by the device is visible on the host
    dest = F.create_fdr(s, 129)
    F.zdtfdr_(dest, ws)
    F.fdrsynchronize(ws)
    F.fdrsynchronize()
    assert np.isfinite(host(dest)).all()
