"""The checks of the reference's network-only integration test
(test/heavytests.jl:17-40: 150 channels of a Voyager-1 spectrogram, rates
0:-0.25:-5; intfdr, fftfdr and zdtfdr peak at the same cell, and fdshift at
the carrier's drift aligns every time step's peak within two channels) on a
synthetic spectrogram of the same shape: Exponential(1) noise plus a carrier
starting between channels 54 and 55 (0-based) and drifting at -2.43
channels per step.  The HDF5 file cannot be fetched here; the peak cell is
the one the FP64 oracle and the integer-shift intfdr agree on (55, 10)
(0-based; test_heavy_analogue_oracle pins it on the CPU).  The GPU test takes
its input the reference's way: the 150 channels are written into a wider
filterbank HDF5 file (Voyager-like header, channels 2000..2149 of 4096) and
read back as data[freq_range, 1, :], with the file's chunks compressed either
by HDF5's deflate + shuffle or by bitshuffle + LZ4 as the Breakthrough Listen
files are (decoded without the plugin the reference loads,
test/heavytests.jl:6)."""
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


def filterbank_file(path, compression="deflate"):
    """The analogue spectrogram inside a 4096-channel filterbank file, its
    chunks deflate + shuffle or bitshuffle + LZ4 compressed."""
    from fdrt_hip import filterbank as FB
    rng = np.random.default_rng(59046)
    wide = rng.exponential(1.0, size=(4096, NT)).astype(np.float32)
    wide[2000:2000 + NF] = spectrogram()
    if compression == "bitshuffle":
        from test_bitshuffle import write_bitshuffled
        write_bitshuffled(path, wide, (NT, 1, 1024), comp=2)
    else:
        FB.write_filterbank(path, wide, {"fch1": 8421.386717353016, "foff": -2.7939677238464355e-06,
                                         "tsamp": 18.253611008, "source_name": "VOYAGER-1"},
                            chunks=(NT, 1, 1024), deflate=4, shuffle=True)
    return range(2000, 2000 + NF)


@pytest.mark.parametrize("compression", ["deflate", "bitshuffle"])
def test_heavy_analogue_ingest(tmp_path, compression):
    from fdrt_hip import filterbank as FB
    freq_range = filterbank_file(tmp_path / "voyager_like.h5", compression)
    s = FB.read_spectrogram(tmp_path / "voyager_like.h5", channels=freq_range)
    assert s.shape == (NF, NT) and np.array_equal(s, spectrogram())
    assert argmax2(Z.intfdr(s, RATES)) == PEAK


@pytest.mark.gpu
@pytest.mark.parametrize("compression", ["deflate", "bitshuffle"])
def test_heavy_analogue_gpu(gpu, tmp_path, compression):
    import fdrt_hip as F
    print(f"heavy-test analogue: 150 x 16, rates 0:-0.25:-5, via a {compression} filterbank HDF5 file", flush=True)
    sys.stdout.flush()
    freq_range = filterbank_file(tmp_path / "voyager_like.h5", compression)
    s = F.read_spectrogram(tmp_path / "voyager_like.h5", channels=freq_range)
    assert np.array_equal(s, spectrogram())
    ws = F.fftfdr_workspace(s)
    assert argmax2(F.fftfdr(ws, RATES).cpu().numpy()) == PEAK
    dedop = F.fdshift(ws, -2.43).cpu().numpy()
    assert set(np.argmax(dedop, axis=0).tolist()) <= {54, 55}
    zws = F.ZDTWorkspace(s, RATES)
    assert argmax2(F.zdtfdr(zws).cpu().numpy()) == PEAK
    # the same channels straight into pinned memory, searched from there
    p = F.pinned_spectrogram(NF, NT)
    F.read_spectrogram(tmp_path / "voyager_like.h5", channels=freq_range, out=p)
    st = F.StreamSearch(NF, NT, (0.0, -0.25, 21))
    hits, thr = st.search([p], snr=5.0)
    st.close()
    assert (PEAK[0] + 1, PEAK[1] + 1) in set(map(tuple, hits[0].tolist()))
