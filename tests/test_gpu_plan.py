"""GPU parity of ragged multi-channel plans (zdt_plan_*): every channel's FDR
against the FP64 oracle, its threshold against fdrstats/fdrdenormalize and
its proto-hits against findprotohits, over several replays of the captured
graph with eager work between them."""
import sys

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
    return float(np.max(np.abs(a.astype(np.float64) - b)) / np.max(np.abs(b)))


def host(t):
    return t.detach().cpu().numpy()


@pytest.fixture(scope="module")
def F(gpu):
    import fdrt_hip
    return fdrt_hip


# (Nf, Nt, r0, dr, Nr): fused shapes, a repeated shape (shared workspace), a
# non-power-of-two Nf and an Nt above the fused limit (generic kernels)
CHANNELS = [
    (1 << 16, 16, -4.0, 0.25, 33),
    (1 << 17, 7, -3.0, 0.5, 13),
    (1 << 16, 16, -4.0, 0.25, 33),
    (24576, 5, -2.0, 0.25, 17),
    (1 << 16, 20, -4.0, 0.25, 33),
]


def make_specs():
    specs = []
    for c, (Nf, Nt, r0, dr, Nr) in enumerate(CHANNELS):
        tones = [(Nf * 0.21, 0.5, 30.0), (Nf * 0.64, -1.25, 40.0), (Nf * 0.9, 2.0, 25.0)]
        specs.append(synth.spectrogram(Nf, Nt, tones, seed=100 + c))
    return specs


def check_channels(F, plan, specs, refs, snr):
    thr_dev = plan.thresholds()
    nh = plan.nhits()
    for c, (Nf, Nt, r0, dr, Nr) in enumerate(CHANNELS):
        fdr = host(plan.fdr(c))
        assert normwise(fdr, refs[c]) <= 2e-6, c
        m, s = H.fdrstats(fdr)
        thr = H.fdrdenormalize(snr, m, s)
        assert abs(thr_dev[c] - thr) <= 1e-9 * abs(thr), (c, thr_dev[c], thr)
        exp = H.findprotohits(fdr, snr, snr=True)
        assert nh[c] == len(exp), c
        np.testing.assert_array_equal(plan.hits(c), exp)


def test_plan_replays_match_oracle(F):
    import torch
    specs = make_specs()
    say("plan: oracle FDRs")
    refs = [Z.fdr64(s, r0, dr, Nr) for s, (_, _, r0, dr, Nr) in zip(specs, CHANNELS)]
    say("plan: create (autotune)")
    plan = F.RaggedPlan(specs, [(r0, dr, Nr) for (_, _, r0, dr, Nr) in CHANNELS], autotune=True)
    infos = [plan.info(c) for c in range(len(CHANNELS))]
    say(f"plan: {infos}")
    assert infos[0]["shape"] == infos[2]["shape"] and infos[0]["shape"] != infos[1]["shape"]
    assert [i["path"] for i in infos] == ["fused", "fused", "fused", "generic", "fused-output"]
    assert infos[0]["n2"] == 512 and infos[1]["n2"] in (512, 1024)
    # the radix cap belongs to the generic path's FFTs; the fused kernels of
    # paths 1 and 2 do not use them and report 0
    assert infos[3]["maxr"] in (8, 16)
    assert infos[0]["maxr"] == infos[1]["maxr"] == infos[4]["maxr"] == 0
    plan.bind(8.0, snr=True, cap=4096)
    for rep in range(3):
        say(f"plan: replay {rep}")
        # eager work between replays: poison the outputs the graph must rewrite
        for d in plan.dests:
            d.fill_(float("nan"))
        plan._hits.fill_(-1)
        plan._nhits.fill_(-1)
        torch.cuda.synchronize()
        plan.launch()
        torch.cuda.synchronize()
        check_channels(F, plan, specs, refs, 8.0)
    plan.close()


def test_plan_absolute_threshold_shared_fdr(F):
    specs = make_specs()
    rates = [(r0, dr, Nr) for (_, _, r0, dr, Nr) in CHANNELS]
    say("plan: absolute thresholds, plan-owned FDR buffers")
    refs = [Z.fdr64(s, *r) for s, r in zip(specs, rates)]
    thr = [float(np.quantile(r, 0.999)) for r in refs]
    plan = F.RaggedPlan(specs, rates, autotune=False)
    plan.bind(thr, snr=False, cap=1 << 14, keep_fdrs=False)
    plan.launch()
    import torch
    torch.cuda.synchronize()
    np.testing.assert_allclose(plan.thresholds(), thr, rtol=0, atol=0)
    for c in range(len(CHANNELS)):
        # the oracle set, minus cells within the FP32 error band of the threshold
        got = {tuple(x) for x in plan.hits(c)}
        ref = refs[c]
        band = 4e-6 * np.max(np.abs(ref))
        sure = {(i + 1, j + 1) for j, i in zip(*np.nonzero(ref.T >= thr[c] + band))}
        maybe = {(i + 1, j + 1) for j, i in zip(*np.nonzero(ref.T >= thr[c] - band))}
        assert sure <= got <= maybe, c
    plan.close()


def test_plan_truncation_and_errors(F):
    specs = make_specs()[:2]
    rates = [(r0, dr, Nr) for (_, _, r0, dr, Nr) in CHANNELS[:2]]
    plan = F.RaggedPlan(specs, rates, autotune=False)
    with pytest.raises(F.ZDTError):
        plan.launch()  # not bound yet
    plan.bind(0.0, snr=True, cap=16)  # threshold at the mean: far more hits than 16
    plan.launch()
    n = plan.nhits()
    assert (n > 16).all()
    with pytest.raises(F.ZDTError):
        plan.hits(0)
    plan.close()
