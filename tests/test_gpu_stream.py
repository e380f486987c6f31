"""GPU parity of the long-integration search (zdt_stream_*): batched rates,
spectrograms streamed from (pinned) host memory, SNR threshold over all
batches, against the FP64 oracle (oracle.zdt_oracle.fdr64) and the oracle's
statistics and findprotohits (src/fdrutils.jl:39-50,104-107,148-169).

The search returns hits and thresholds, not FDRs, so the checks are:
  * threshold: the oracle's thr = snr*s + m over exactly the requested Nr
    columns (relative 1e-5: the FDR itself is within 2e-6 normwise);
  * hits: identical to findall(>= thr) on the FP64 FDR at the search's own
    threshold, except for cells within eps of it, eps = 4 ulp(thr) plus the
    2e-6 normwise FDR tolerance.
"""
import sys
from fractions import Fraction

import numpy as np
import pytest

from oracle import host_oracle as H
from oracle import synth
from oracle import zdt_oracle as Z

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


def check_hits(hits, ref64, thr, cols=None, tol=2e-6):
    """hits (1-based (i, j)) against findall(ref64 >= thr); cols: the 0-based
    rate columns ref64 holds (None: all of them, in order); tol: the FDR's
    normwise tolerance."""
    scale = float(np.max(np.abs(ref64)))
    eps = 4 * float(np.spacing(np.float32(thr))) + tol * scale
    if cols is None:
        cols = list(range(ref64.shape[1]))
    colset = {c + 1: q for q, c in enumerate(cols)}
    g = {(int(i), int(j)) for i, j in hits.tolist() if int(j) in colset}
    r = set()
    band = set()
    for c, q in colset.items():
        x = ref64[:, q]
        r |= {(int(i) + 1, c) for i in np.nonzero(x >= thr)[0]}
        band |= {(int(i) + 1, c) for i in np.nonzero(np.abs(x - thr) <= eps)[0]}
    assert (g ^ r) <= band, sorted(g ^ r)[:10]
    return len(r)


def batch_hits(F, spec, r0, dr, Nr, Nrb, thr, precision="fp32"):
    """The proto-hits a batched search must return, by a full read of every
    batch FDR: each batch is a workspace of Nrb rates from r0 + b*Nrb*dr (the
    search's own batch, so the same FDR), its requested columns are read by
    the plain findprotohits, and the columns are offset by b*Nrb."""
    out = []
    for b in range((Nr + Nrb - 1) // Nrb):
        ws = F.ZDTWorkspace(spec, r0 + b * Nrb * dr, dr, Nrb, precision=precision)
        fdr = F.zdtfdr(ws)
        h = F.findprotohits(fdr[:, :min(Nrb, Nr - b * Nrb)], thr)
        h[:, 1] += b * Nrb
        out.append(h)
        ws.close()
    return np.concatenate(out)


def oracle_thr(ref64, snr):
    m, s = H.fdrstats(ref64)
    return H.fdrdenormalize(snr, m, s)


@pytest.mark.parametrize("mode", ["resident", "recompute"])
def test_stream_search_vs_fp64(F, mode):
    """Nr = 33 in batches of 10: four batches, the last holding 3 requested
    rates and 7 extra ones that must not enter the statistics or the hits."""
    import torch
    specs = specs3()
    # 36 MiB: one 2^16 x 10 FDR and the workspace fit (34.6 MB), all four
    # batch FDRs (42.5 MB) do not
    budget = 0 if mode == "resident" else 36 << 20
    st = F.StreamSearch(NF, NT, (R0, DR, NR), Nrb=10, hbm_bytes=budget)
    say(f"stream {mode}: Nrb={st.Nrb} nbatch={st.nbatch} resident={st.resident}")
    assert (st.Nrb, st.nbatch) == (10, 4) and st.resident == (mode == "resident")
    # one pinned spectrogram, one pageable tensor, one numpy array
    p = F.pinned_spectrogram(NF, NT)
    p.copy_(torch.from_numpy(np.asfortranarray(specs[0])))
    inputs = [p, torch.from_numpy(specs[1].T.copy()).t(), specs[2]]
    hits, thr = st.search(inputs, snr=8.0, cap=1 << 16)
    for i, s in enumerate(specs):
        ref = Z.fdr64(s, R0, DR, NR)
        t = oracle_thr(ref, 8.0)
        assert abs(thr[i] - t) <= 1e-5 * abs(t), (i, thr[i], t)
        assert hits[i][:, 1].max() <= NR
        assert check_hits(hits[i], ref, thr[i]) > 0
    st.close()


@pytest.mark.parametrize("di,dj", [(1, 1), (2, 0)])
def test_stream_findhits(F, di, dj):
    """zdt_stream_findhits: the search ending in hits.  Nr = 33 in batches of
    10, resident: the hits of each spectrogram equal, exactly, the clustering
    of the stream's own proto-hits over the hcat of the batch FDRs (the
    search's own batches, rebuilt by workspaces), on the device and by the
    oracle on the host.  A non-resident stream is refused, and too small a
    proto-hit staging is reported."""
    from fdrt_hip import _lib
    from oracle import cluster_oracle as CO
    specs = specs3()
    st = F.StreamSearch(NF, NT, (R0, DR, NR), Nrb=10)
    assert st.resident and st.nbatch == 4
    snr = 6.0
    got, nph, thr = st.findhits(specs, snr, di=di, dj=dj)
    ph, thr2 = st.search(specs, snr)
    assert np.array_equal(thr, thr2)
    for i, spec in enumerate(specs):
        say(f"stream findhits {i}: {nph[i]} proto-hits, {len(got[i].size)} hits")
        assert nph[i] == len(ph[i]) > 0
        fdrs = []
        for b in range(4):
            ws = F.ZDTWorkspace(spec, R0 + b * 10 * DR, DR, 10)
            fdrs.append(F.zdtfdr(ws))
            ws.close()
        ref = F.clusterprotohits(fdrs, ph[i], di, dj)
        for name in ("peak", "value", "size", "box"):
            assert np.array_equal(getattr(got[i], name), getattr(ref, name)), name
        if i == 0:
            o = CO.clusterprotohits([f.cpu().numpy() for f in fdrs], ph[i], di, dj)
            for name in ("peak", "value", "size", "box"):
                assert np.array_equal(getattr(got[i], name), o[name]), name
    # too few hits or proto-hits staged: reported, not wrong
    with pytest.raises(_lib.ZDTError, match="capacity"):
        st.findhits(specs[:1], snr, di=di, dj=dj, cap=1)
    with pytest.raises(_lib.ZDTError, match="proto-hits"):
        st.findhits(specs[:1], snr, di=di, dj=dj, pcap=3)
    st.close()
    rec = F.StreamSearch(NF, NT, (R0, DR, NR), Nrb=10, hbm_bytes=36 << 20)
    assert not rec.resident
    with pytest.raises(_lib.ZDTError, match="resident"):
        rec.findhits(specs[:1], snr)
    rec.close()


def test_stream_auto_batch_and_truncation(F):
    specs = specs3()[:1]
    st = F.StreamSearch(NF, NT, (R0, DR, NR))
    assert st.Nrb == NR and st.nbatch == 1 and st.resident
    hits, thr = st.search(specs, snr=8.0)
    ref = Z.fdr64(specs[0], R0, DR, NR)
    assert abs(thr[0] - oracle_thr(ref, 8.0)) <= 1e-5 * abs(thr[0])
    check_hits(hits[0], ref, thr[0])
    with pytest.raises(F.ZDTError):
        st.search(specs, snr=0.0, cap=8)
    with pytest.raises(ValueError):
        st.search([np.zeros((NF, NT + 1), np.float32)])
    st.close()


@pytest.mark.parametrize("Nf,Nt", [(1 << 17, 20), (1 << 21, 20), (24576, 9)])
def test_stream_fused_output_and_generic_paths(F, Nf, Nt):
    """The statistics come from the output kernel (fused output, Nt > 16;
    at 2^21 its column DFTs have N1 = 1024, two rows per hit tile) or from
    colstats over each batch (generic path, Nf not a power of two)."""
    r0, dr, Nr, Nrb = -2.0, 0.125, 33, 11
    specs = [synth.spectrogram(Nf, Nt, [(Nf * 0.3, 1.5 - i, 30.0), (Nf * 0.7, -1.0, 25.0)], seed=400 + i)
             for i in range(2)]
    st = F.StreamSearch(Nf, Nt, (r0, dr, Nr), Nrb=Nrb)
    say(f"stream Nf={Nf} Nt={Nt}: Nrb={st.Nrb} nbatch={st.nbatch}")
    assert (st.Nrb, st.nbatch) == (Nrb, 3)
    hits, thr = st.search(specs, snr=7.0, cap=1 << 16)
    for i, s in enumerate(specs):
        ref = Z.fdr64(s, r0, dr, Nr)
        assert abs(thr[i] - oracle_thr(ref, 7.0)) <= 1e-5 * abs(thr[i])
        check_hits(hits[i], ref, thr[i])
        # exactly the hits of a full read of the same batch FDRs (the fused
        # paths count only the tiles their run maxima allow)
        assert np.array_equal(hits[i], batch_hits(F, s, r0, dr, Nr, Nrb, thr[i]))
    st.close()


@pytest.mark.parametrize("precision", ["split-bf16", "split-fp16"])
@pytest.mark.parametrize("Nf,Nt", [(1 << 17, 16), (1 << 17, 20)])
def test_stream_split_precision(F, Nf, Nt, precision):
    """The search with its column DFTs on the matrix cores (fused path at
    Nt <= 16, fused output above): thresholds within 1e-4 of the oracle's,
    hits the FP64 oracle's up to cells within the split modes' 1e-4 normwise
    tolerance of the threshold."""
    r0, dr, Nr, Nrb = -2.0, 0.125, 33, 11
    specs = [synth.spectrogram(Nf, Nt, [(Nf * 0.3, 1.5 - i, 30.0), (Nf * 0.7, -1.0, 25.0)], seed=500 + i)
             for i in range(2)]
    st = F.StreamSearch(Nf, Nt, (r0, dr, Nr), Nrb=Nrb, precision=precision)
    say(f"stream {precision} Nf={Nf} Nt={Nt}: Nrb={st.Nrb} nbatch={st.nbatch}")
    assert st.precision == precision
    hits, thr = st.search(specs, snr=7.0, cap=1 << 16)
    for i, s in enumerate(specs):
        ref = Z.fdr64(s, r0, dr, Nr)
        assert abs(thr[i] - oracle_thr(ref, 7.0)) <= 1e-4 * abs(thr[i])
        assert check_hits(hits[i], ref, thr[i], tol=1e-4) > 0
        assert np.array_equal(hits[i], batch_hits(F, s, r0, dr, Nr, Nrb, thr[i], precision))
    st.close()
    with pytest.raises(ValueError):
        F.StreamSearch(Nf, Nt, (r0, dr, Nr), precision="bf16")


def test_stream_search_files(F, tmp_path):
    """search_files: the same hits and thresholds as search() over the
    arrays, with the spectrograms read on the reader thread, two files at a
    time, from filterbank files (deflate, and bitshuffle/LZ4 decoded without
    the plugin) as a channel window of a wider file."""
    from fdrt_hip import filterbank as FB
    from test_bitshuffle import write_bitshuffled
    specs = specs3()
    paths = []
    for i, s in enumerate(specs):
        wide = np.random.default_rng(i).exponential(1.0, size=(NF + 100, NT)).astype(np.float32)
        wide[50:50 + NF] = s
        path = tmp_path / f"spec{i}.h5"
        if i == 0:
            FB.write_filterbank(path, wide, chunks=(NT, 1, 8192), deflate=1)
        else:
            write_bitshuffled(path, wide, (NT, 1, 8192))
        paths.append(path)
    st = F.StreamSearch(NF, NT, (R0, DR, NR), Nrb=10)
    h0, t0 = st.search(specs, snr=8.0)
    h1, t1 = st.search_files(paths, snr=8.0, channels=range(50, 50 + NF), group=2)
    st.close()
    assert len(h1) == 3 and np.array_equal(t0, t1)
    assert all(np.array_equal(a, b) for a, b in zip(h0, h1))


def test_stream_profile(F):
    """profile=True: every kernel of a search is timed on the search's
    workspace, the hit extraction included; the same thresholds and hits as
    an unprofiled search."""
    specs = specs3()[:2]
    st = F.StreamSearch(NF, NT, (R0, DR, NR), Nrb=10, profile=True)
    assert st.path == "fused"
    h0, t0 = st.search(specs, snr=8.0)
    st.reset_profile()
    h1, t1 = st.search(specs, snr=8.0)
    kt = st.kernel_times()
    st.close()
    names = {k for k, _, _ in kt}
    say(f"stream profile: {[(k, round(ms, 3), n) for k, ms, n in kt]}")
    # the fused path's k3 writes tile maxima, so the count reads only the
    # tiles that can hold a proto-hit (hits_count_sparse)
    assert {"k1_input", "k2_czt", "k3_output", "hits_threshold", "hits_count_sparse", "hits_write"} <= names
    launches = dict((k, n) for k, _, n in kt)
    assert launches["k1_input"] == 2 and launches["k3_output"] == 2 * 4 and launches["hits_count_sparse"] == 2 * 4
    assert all(ms > 0 for _, ms, _ in kt)
    assert np.array_equal(t0, t1) and all(np.array_equal(a, b) for a, b in zip(h0, h1))
    plain = F.StreamSearch(NF, NT, (R0, DR, NR), Nrb=10)
    h2, t2 = plain.search(specs, snr=8.0)
    plain.close()
    assert np.array_equal(t1, t2) and all(np.array_equal(a, b) for a, b in zip(h1, h2))
    with pytest.raises(F.ZDTError):
        F.StreamSearch(NF, NT, (R0, DR, NR), Nrb=10).reset_profile()


def long_spectrogram(seed):
    """One long-integration spectrogram: 2^22 channels x 64 steps,
    Exponential(1) noise + 64 tones, amplitudes log-uniform 5..100, rates
    uniform in [-8, 8]."""
    Nf, Nt = 1 << 22, 64
    rng = np.random.default_rng(seed)
    s = np.asfortranarray(rng.exponential(1.0, size=(Nt, Nf)).astype(np.float32).T)
    c0, rate = rng.uniform(0, Nf, 64), rng.uniform(-8, 8, 64)
    amp = np.exp(rng.uniform(np.log(5), np.log(100), 64)).astype(np.float32)
    for t in range(Nt):
        np.add.at(s[:, t], np.round(c0 + rate * t).astype(np.int64) % Nf, amp)
    return s, rate


@pytest.mark.timeout(420)
def test_long_integration_shape_vs_fp64(F):
    """2^22 x 64 with rates -8:1/64:8 (Nr = 1025): the search resident and
    with a budget that forces recomputed batches give the same thresholds and
    hits; sampled columns of the transform (a workspace of the same shape)
    are within 2e-6 of fdr64, and the search's hits in those columns are the
    FP64 oracle's at its threshold."""
    import torch
    Nf, Nt, r0, dr, Nr = 1 << 22, 64, -8.0, 1.0 / 64, 1025
    say("long 2^22 x 64 x 1025: building spectrogram")
    s, tone_rates = long_spectrogram(2024)
    p = F.pinned_spectrogram(Nf, Nt)
    p.copy_(torch.from_numpy(s))
    res = F.StreamSearch(Nf, Nt, (r0, dr, Nr))
    say(f"long resident: Nrb={res.Nrb} nbatch={res.nbatch} resident={res.resident}")
    assert res.resident
    h1, t1 = res.search([p], snr=10.0)
    # the same search ending in hits (4.3e9 cells: the general clustering
    # kernels, 64-bit cell indices), checked against the FDR below
    hh, nph, th = res.findhits([p], snr=10.0, cap=1 << 18)
    assert nph[0] == len(h1[0]) and th[0] == t1[0]
    res.close()
    torch.cuda.empty_cache()
    # 16 GiB: room for the workspace (~8 GB at 205 rates), the two
    # spectrogram buffers and one 205-rate FDR (3.4 GB), not for five
    rec = F.StreamSearch(Nf, Nt, (r0, dr, Nr), Nrb=205, hbm_bytes=16 << 30)
    say(f"long recompute: Nrb={rec.Nrb} nbatch={rec.nbatch} resident={rec.resident}")
    assert not rec.resident and rec.nbatch == 5
    h2, t2 = rec.search([p], snr=10.0)
    rec.close()
    torch.cuda.empty_cache()
    # the two batchings pair their rates around different centres, so the
    # FDRs agree to rounding only: thresholds to 1e-6, hit sets up to cells
    # at the threshold (checked against the oracle below)
    assert abs(t1[0] - t2[0]) <= 1e-6 * abs(t1[0])
    d12 = set(map(tuple, h1[0].tolist())) ^ set(map(tuple, h2[0].tolist()))
    assert len(d12) <= max(4, len(h1[0]) // 1000)
    # sampled columns (the FP64 oracle takes ~12 s per column at 2^22 x 64):
    # the lowest rate, one tone's rate and the highest rate
    cols = sorted({0, int(round((tone_rates[0] - r0) / dr)), 1024})
    say(f"long fdr64 columns {cols}")
    rates = [Fraction(-8) + Fraction(j, 64) for j in cols]
    ref = Z.fdr64(s, None, rates=rates)
    # the transform itself: the search's resident configuration (one
    # workspace, all 1025 rates, a 17 GB FDR), sampled columns
    ws = F.ZDTWorkspace(p.to("cuda"), r0, dr, Nr)
    assert ws.path == "fused-output"
    fdr = F.zdtfdr(ws)
    out = fdr[:, cols].cpu().numpy()
    # the resident search computed this same FDR (one batch of all 1025
    # rates); its hit count skipped tiles by k3's run maxima, the plain
    # findprotohits reads every cell: the hit lists are identical
    full = F.findprotohits(fdr, t1[0])
    say(f"long: {len(full)} proto-hits by a full read, {len(h1[0])} by the search")
    assert np.array_equal(full, h1[0])
    ref_hits = F.clusterprotohits(fdr, full, 1, 1, capacity=1 << 18)
    say(f"long: {len(ref_hits.size)} hits")
    for name in ("peak", "value", "size", "box"):
        assert np.array_equal(getattr(hh[0], name), getattr(ref_hits, name)), name
    del fdr
    ws.close()
    torch.cuda.empty_cache()
    for q, j in enumerate(cols):
        err = float(np.max(np.abs(out[:, q] - ref[:, q])) / np.max(np.abs(ref[:, q])))
        say(f"long column {j}: normwise {err:.2e}")
        assert err <= 2e-6
    check_hits(h1[0], ref, t1[0], cols=cols)
    check_hits(h2[0], ref, t2[0], cols=cols)
    # the column mean is the spectrogram's mean power summed over time
    # (every drift line visits each step once): m = sum(s)/Nf, and s > 0
    m_exact = float(s.astype(np.float64).sum()) / Nf
    assert t1[0] > m_exact
