"""GPU parity of the split-precision modes (zdt_create's precision argument):
the 1024-point column DFTs of input! and output! (k1, k3; the frequency-axis
transforms of src/zdtfdr.jl:119-120,211,330) as BF16 or FP16 matrix-core
products of split hi+lo operands, FP32 accumulation.  The chirp-Z over time,
its FP64 phases and the inner DFT of k2 stay FP32.

Bars: normwise max|FDR - fdr64| / max|fdr64| <= 1e-4 against the FP64
oracle; the proto-hit set at SNR 8 on the headline workload (2^20 x 16 x 129)
equal to the FP64 oracle's outside the tie band |fdr64 - thr64| <= eps with
eps = 4 ulp(thr64) + 1e-4 max|fdr64| + |thr - thr64|.
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

SPLIT = ["split-bf16", "split-fp16"]
TOL = 1e-4


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def normwise(a, b):
    return float(np.max(np.abs(np.asarray(a, np.float64) - b)) / np.max(np.abs(b)))


def host(t):
    return t.detach().cpu().numpy()


@pytest.fixture(scope="module")
def F(gpu):
    import fdrt_hip
    return fdrt_hip


@pytest.mark.parametrize("prec", SPLIT)
@pytest.mark.parametrize("Nf,Nt,r0,dr,Nr,path", [
    (1 << 17, 16, -4.0, 0.25, 33, "fused"),
    (1 << 18, 9, -2.0, 0.125, 17, "fused"),
    (1 << 20, 16, -3.0, 0.1875, 33, "fused"),
    (1 << 18, 32, -1.0, 1.0 / 32, 65, "fused-output"),
    (1 << 22, 20, -2.0, 0.25, 5, "fused-output"),
])
def test_split_precision_vs_fp64(F, prec, Nf, Nt, r0, dr, Nr, path):
    s = synth.spectrogram(Nf, Nt, [(Nf * 0.31, 1.5, 25.0), (Nf * 0.77, -2.75, 40.0)], seed=Nf % 1000 + Nt)
    say(f"{prec} Nf={Nf} Nt={Nt} Nr={Nr}")
    ws = F.ZDTWorkspace(s, r0, dr, Nr, precision=prec)
    assert (ws.path, ws.precision) == (path, prec)
    out = host(F.zdtfdr(ws))
    ref = Z.fdr64(s, r0, dr, Nr)
    err = normwise(out, ref)
    # the FP32 path on the same input: the split path must run different
    # arithmetic (a nonzero difference) within the split bar
    w32 = F.ZDTWorkspace(s, r0, dr, Nr)
    out32 = host(F.zdtfdr(w32))
    d = normwise(out, out32.astype(np.float64))
    say(f"{prec} Nf={Nf}: normwise vs fp64 {err:.2e} (fp32 {normwise(out32, ref):.2e}), vs fp32 {d:.2e}")
    assert err <= TOL
    assert d > 0.0


@pytest.mark.parametrize("prec", SPLIT)
def test_split_precision_step_functions_and_batches(F, prec):
    """The step API and r0 batches run the same k1/k3 variants: input!,
    zdtfdr! with no destination then output!, and a second r0."""
    Nf, Nt, dr, Nr = 1 << 17, 12, 0.25, 21
    s = synth.spectrogram(Nf, Nt, [(Nf * 0.4, 0.75, 30.0)], seed=77)
    ws = F.ZDTWorkspace(s, -2.5, dr, Nr, precision=prec)
    for r0 in (-2.5, 1.0):
        F.zdtfdr_(ws, r0=r0)
        dest = F.create_fdr(s, Nr)
        F.output_(dest, ws)
        assert normwise(host(dest), Z.fdr64(s, r0, dr, Nr)) <= TOL


def test_split_precision_unsupported(F):
    s = synth.spectrogram(1 << 16, 8, [(1000.0, 1.0, 20.0)], seed=5)
    for prec in SPLIT:
        with pytest.raises(F.ZDTError):  # N2 = 512 at 2^16 channels: no split variant
            F.ZDTWorkspace(s, -1.0, 0.25, 9, precision=prec)
        with pytest.raises(F.ZDTError):  # the generic path has none either
            F.ZDTWorkspace(synth.spectrogram(6000, 4, [], seed=1), -1.0, 0.25, 9, precision=prec)
    with pytest.raises(ValueError):
        F.ZDTWorkspace(s, -1.0, 0.25, 9, precision="fp8")


CFG2 = json.loads((Path(__file__).parent / "golden" / "cfg2_hits_snr8.json").read_text())


@pytest.mark.parametrize("prec", SPLIT)
def test_split_precision_headline_hit_set(F, prec):
    """2^20 x 16 x 129, SNR 8: sampled columns within 1e-4 of fdr64 and the
    proto-hit set against the FP64 oracle's (tests/golden/cfg2_hits_snr8.json)."""
    from oracle.synth import random_tones
    s = synth.cfg2(seed=1234)
    tones = random_tones(1 << 20, 16, (10, 50), (-4, 4), 1234)
    say(f"{prec} cfg2 2^20 x 16 x 129")
    ws = F.ZDTWorkspace(s, -4.0, 0.0625, 129, precision=prec)
    assert (ws.path, ws.precision) == ("fused", prec)
    fdr = F.zdtfdr(ws)
    near = {min(128, max(0, int(np.floor((r + 4.0) * 16)) + d)) for _, r, _ in tones for d in (0, 1)}
    cols = sorted({0, 64, 128} | near)
    ref = Z.fdr64(s, None, rates=[Fraction(-4) + Fraction(j, 16) for j in cols])
    err = normwise(host(fdr[:, cols]), ref)
    say(f"{prec} cfg2 {len(cols)} columns: normwise {err:.2e}")
    assert err <= TOL
    hits = F.findprotohits(fdr, 8, snr=True)
    thr = F.fdrdenormalize(8, fdr)
    thr64 = CFG2["thr64"]
    assert abs(thr - thr64) <= 1e-5 * thr64
    eps = 4 * float(np.spacing(np.float32(thr64))) + TOL * CFG2["max_abs_fdr64"] + abs(thr - thr64)
    assert eps <= CFG2["band_abs"]
    band = {(i, j) for i, j, dd in CFG2["band"] if abs(dd) <= eps}
    g = set(map(tuple, hits.tolist()))
    r = set(map(tuple, CFG2["hits"]))
    say(f"{prec} cfg2 hits: {len(g)} (FP64 {len(r)}), symmetric difference {len(g ^ r)}, band {len(band)}")
    assert (g ^ r) <= band
