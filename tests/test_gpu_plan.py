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
    # the autotuner picks each shape's path among the ones its shape allows:
    # fused (power-of-two Nf <= 2^20, Nt <= 16), fused output (power-of-two
    # Nf), generic (any)
    allowed = [{"fused", "fused-output", "generic"}, {"fused", "fused-output", "generic"},
               {"fused", "fused-output", "generic"}, {"generic"}, {"fused-output", "generic"}]
    for i, ok in zip(infos, allowed):
        assert i["path"] in ok, infos
        # the radix cap belongs to the generic path's FFTs; the fused kernels
        # of paths 1 and 2 do not use them and report 0
        assert i["maxr"] in ((8, 16) if i["path"] == "generic" else (0,)), infos
        if i["path"] == "fused":
            assert i["n2"] == 512 or (i["n2"] == 1024 and i is infos[1]), infos
    assert all(plan.tuning(c)["source"] == "timed" for c in range(len(CHANNELS)))
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


def test_plan_tuning_table(F, tmp_path):
    """The autotuner's choices persist in a tuning table: a second plan of the
    same shapes reads them instead of timing, makes the same choices, and
    its results are the same; lines that do not parse are ignored."""
    import torch
    specs = make_specs()
    rates = [(r0, dr, Nr) for (_, _, r0, dr, Nr) in CHANNELS]
    table = tmp_path / "tune.txt"
    table.write_text("not a tuning line\nzdt-tune-v1 gfx000 1 2 3\n")
    F.set_tune_table(str(table))
    try:
        p1 = F.RaggedPlan(specs, rates, autotune=True)
        lines = [ln for ln in table.read_text().splitlines() if ln.startswith("zdt-tune-v1 ") and len(ln.split()) == 11]
        assert len(lines) == 4  # one per distinct shape
        assert all(p1.tuning(c)["source"] == "timed" for c in range(len(CHANNELS)))
        p2 = F.RaggedPlan(specs, rates, autotune=True)
        for c in range(len(CHANNELS)):
            assert p2.tuning(c)["source"] == "table", c
            assert p2.info(c) == p1.info(c), c
            assert p2.tuning(c)["ms"] == pytest.approx(p1.tuning(c)["ms"], rel=1e-3)
        for p in (p1, p2):
            p.bind(8.0, snr=True, cap=4096)
            p.launch()
        torch.cuda.synchronize()
        for c in range(len(CHANNELS)):
            np.testing.assert_array_equal(host(p1.fdr(c)), host(p2.fdr(c)))
            np.testing.assert_array_equal(p1.hits(c), p2.hits(c))
        p1.close()
        p2.close()
    finally:
        F.set_tune_table(None)
    # with no table, nothing is read
    p3 = F.RaggedPlan(specs[:1], rates[:1], autotune=True)
    assert p3.tuning(0)["source"] == "timed"
    p3.close()

