"""The ragged multi-channel workload end to end: 64 coarse channels (Nf from
{2^18, 2^19, 2^20}, Nt from {16, 32, 60}, 8 tones each, rates -4:1/Nt:4, the
channels of bench.py's ragged mode, seed 4321) through one RaggedPlan graph
launch with autotuning, SNR-10 proto-hits compacted on the device.  It stands
in for the reference user's loop over channels around zdtfdr!
(src/zdtfdr.jl:349-421), fdrstats/fdrdenormalize (src/fdrutils.jl:39-50,
104-119) and findprotohits (:148-169).

Checks, per channel: every column sums to the spectrogram's total power (each
drift line visits every time step once); the plan's threshold equals
10 s + m from FP64 statistics of its FDR; the plan's hit list is strictly
increasing in column-major order, every listed cell is >= the threshold and
the list is as long as the count of such cells (so it is exactly
findall(>=(thr))).  Per distinct shape: 4 rate columns against the FP64
oracle (normwise <= 2e-6).
"""
import sys
from fractions import Fraction

import numpy as np
import pytest

from oracle import zdt_oracle as Z

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def F(gpu):
    import fdrt_hip
    return fdrt_hip


@pytest.mark.timeout(600)
def test_ragged_64_channels(F):
    import torch

    import bench
    dev = torch.device("cuda", 0)
    shapes, specs, rates = bench.ragged_channels(64, 4321, dev)
    say(f"ragged: {len(set(shapes))} shapes over 64 channels")
    plan = F.RaggedPlan(specs, rates, autotune=True)
    plan.bind(10.0, snr=True, cap=1 << 16, keep_fdrs=True)
    for rep in range(2):
        for d in plan.dests:
            d.fill_(float("nan"))
        torch.cuda.synchronize()
        plan.launch()
        torch.cuda.synchronize()
    thr = plan.thresholds()
    nh = plan.nhits()
    seen = set()
    for c, (Nf, Nt) in enumerate(shapes):
        fdr = plan.fdr(c)
        Nr = rates[c][2]
        colsum = fdr.double().sum(dim=0)
        total = specs[c].double().sum()
        assert torch.allclose(colsum, total.expand_as(colsum), rtol=1e-5), c
        f64 = fdr.double()
        m = float(f64[:, 0].mean())
        s = float(f64.std(dim=0).min())
        assert abs(thr[c] - (10.0 * s + m)) <= 1e-9 * abs(thr[c]), (c, thr[c], 10.0 * s + m)
        del f64
        hits = plan.hits(c)
        assert len(hits) == nh[c] and nh[c] > 0, c
        lin = torch.from_numpy((hits[:, 0] - 1) + Nf * (hits[:, 1] - 1)).to(dev)
        flat = fdr.t().reshape(-1)  # column-major linear order
        assert bool((lin[1:] > lin[:-1]).all()), c
        assert bool((flat[lin].double() >= thr[c]).all()), c
        assert int((flat.double() >= thr[c]).sum()) == nh[c], c
        if (Nf, Nt) not in seen:
            seen.add((Nf, Nt))
            cols = [0, Nr // 3, Nr // 2, Nr - 1]
            # the workspace holds dr as Float32 (1/60 is not exact)
            rs = [Fraction(-4) + Fraction(float(np.float32(1.0 / Nt))) * j for j in cols]
            ref = Z.fdr64(specs[c].cpu().numpy(), None, rates=rs)
            out = fdr[:, cols].cpu().numpy().astype(np.float64)
            err = float(np.max(np.abs(out - ref)) / np.max(np.abs(ref)))
            say(f"ragged channel {c} ({Nf} x {Nt}, {plan.info(c)['path']}): {nh[c]} hits, columns {cols} "
                f"normwise {err:.2e}")
            assert err <= 2e-6, c
    assert len(seen) >= 8
    plan.close()
