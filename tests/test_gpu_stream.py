"""GPU parity of the long-integration search (zdt_stream_*): batched rates,
spectrograms streamed from (pinned) host memory, SNR threshold over all
batches, against the same computation through the workspace API and the
oracle's multi-batch findprotohits."""
import sys

import numpy as np
import pytest

from oracle import host_oracle as H
from oracle import synth

pytestmark = pytest.mark.gpu

NF, NT, R0, DR, NR = 1 << 16, 16, -4.0, 0.25, 33


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def F(gpu):
    import fdrt_hip
    return fdrt_hip


def specs3():
    out = []
    for i in range(3):
        tones = [(NF * (0.15 + 0.3 * i), 1.75 - i, 25.0), (NF * 0.55, -3.5 + i, 20.0)]
        out.append(synth.spectrogram(NF, NT, tones, seed=300 + i))
    return out


def expected(F, spec, Nrb, nb, snr):
    """The same search through the workspace API + the oracle's findprotohits."""
    ws = F.ZDTWorkspace(spec, R0, DR, Nrb)
    r0s = [R0 + b * Nrb * DR for b in range(nb)]
    dests = [F.create_fdr(spec, Nrb) for _ in r0s]
    F.zdtfdr_(dests, ws, r0=r0s)
    fdrs = [d.cpu().numpy() for d in dests]
    m, s = H.fdrstats(fdrs)
    return H.findprotohits(fdrs, snr, snr=True), H.fdrdenormalize(snr, m, s)


@pytest.mark.parametrize("mode", ["resident", "recompute"])
def test_stream_search_matches_batched_zdt(F, mode):
    import torch
    specs = specs3()
    budget = 0 if mode == "resident" else 30 << 20
    st = F.StreamSearch(NF, NT, (R0, DR, NR), Nrb=10, hbm_bytes=budget)
    say(f"stream {mode}: Nrb={st.Nrb} nbatch={st.nbatch} resident={st.resident}")
    assert (st.Nrb, st.nbatch) == (10, 4) and st.resident == (mode == "resident")
    # one pinned spectrogram, one pageable tensor, one numpy array
    p = F.pinned_spectrogram(NF, NT)
    p.copy_(torch.from_numpy(np.asfortranarray(specs[0])))
    inputs = [p, torch.from_numpy(specs[1].T.copy()).t(), specs[2]]
    hits, thr = st.search(inputs, snr=8.0, cap=1 << 16)
    for i, s in enumerate(specs):
        exp, t = expected(F, s, 10, 4, 8.0)
        assert abs(thr[i] - t) <= 1e-9 * abs(t), (i, thr[i], t)
        np.testing.assert_array_equal(hits[i], exp)
    st.close()


def test_stream_auto_batch_and_truncation(F):
    specs = specs3()[:1]
    st = F.StreamSearch(NF, NT, (R0, DR, NR))
    assert st.Nrb == NR and st.nbatch == 1 and st.resident
    hits, thr = st.search(specs, snr=8.0)
    exp, t = expected(F, specs[0], NR, 1, 8.0)
    np.testing.assert_array_equal(hits[0], exp)
    with pytest.raises(F.ZDTError):
        st.search(specs, snr=0.0, cap=8)
    with pytest.raises(ValueError):
        st.search([np.zeros((NF, NT + 1), np.float32)])
    st.close()


@pytest.mark.parametrize("Nf,Nt", [(1 << 17, 20), (24576, 9)])
def test_stream_fused_output_and_generic_paths(F, Nf, Nt):
    """The statistics come from the output kernel (fused output, Nt > 16) or
    from colstats over each batch (generic path, Nf not a power of two)."""
    r0, dr, Nr, Nrb = -2.0, 0.125, 33, 11
    specs = [synth.spectrogram(Nf, Nt, [(Nf * 0.3, 1.5 - i, 30.0), (Nf * 0.7, -1.0, 25.0)], seed=400 + i)
             for i in range(2)]
    st = F.StreamSearch(Nf, Nt, (r0, dr, Nr), Nrb=Nrb)
    say(f"stream Nf={Nf} Nt={Nt}: Nrb={st.Nrb} nbatch={st.nbatch}")
    assert (st.Nrb, st.nbatch) == (Nrb, 3)
    hits, thr = st.search(specs, snr=7.0, cap=1 << 16)
    for i, s in enumerate(specs):
        ws = F.ZDTWorkspace(s, r0, dr, Nrb)
        r0s = [r0 + b * Nrb * dr for b in range(3)]
        dests = [F.create_fdr(s, Nrb) for _ in r0s]
        F.zdtfdr_(dests, ws, r0=r0s)
        fdrs = [d.cpu().numpy() for d in dests]
        m, sd = H.fdrstats(fdrs)
        t = H.fdrdenormalize(7.0, m, sd)
        assert abs(thr[i] - t) <= 1e-9 * abs(t), (i, thr[i], t)
        np.testing.assert_array_equal(hits[i], H.findprotohits(fdrs, t))
    st.close()
