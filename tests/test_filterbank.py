"""Filterbank HDF5 ingest (fdrt_hip.filterbank): the reference's integration
test reads its spectrogram as h5["data"][freq_range, 1, :] from a
Breakthrough Listen filterbank file (test/heavytests.jl:10-14), dataset
dimensions (time, if, chan) in HDF5's C order.  Files are written here with
the same library (libhdf5) and checked back against the numpy arrays they
were written from: hyperslabs, a second IF, chunked deflate/shuffle storage,
header attributes, the frequency axis, reads into caller buffers, and the
errors (an unavailable filter, bad ranges and buffers, a missing dataset)."""
import numpy as np
import pytest
import torch

from fdrt_hip import filterbank as FB

HDR = {"fch1": 8421.386717353016, "foff": -2.7939677238464355e-06, "tsamp": 18.253611008,
       "tstart": 59046.92634259259, "source_name": "VOYAGER-1", "telescope_id": 6}


def spec(nf=1000, nt=16, seed=1):
    return np.random.default_rng(seed).exponential(1.0, size=(nf, nt)).astype(np.float32)


def test_roundtrip_and_header(tmp_path):
    a = spec()
    FB.write_filterbank(tmp_path / "a.h5", a, HDR)
    with FB.FilterbankFile(tmp_path / "a.h5") as f:
        assert f.shape == (1000, 1, 16) and f.filters == []
        assert f.header["source_name"] == "VOYAGER-1" and f.header["telescope_id"] == 6
        assert f.header["fch1"] == HDR["fch1"] and f.header["nchans"] == 1000
        x = f.read()
        assert x.flags["F_CONTIGUOUS"] and x.shape == (1000, 16)
        assert np.array_equal(x, a)
        # data[freq_range, 1, :] with freq_range = 101:250 (1-based) and a time window
        assert np.array_equal(f.read(channels=range(100, 250)), a[100:250])
        assert np.array_equal(f.read(channels=slice(999, 1000), times=range(3, 11)), a[999:1000, 3:11])
        fr = f.frequencies(range(10, 13))
        assert np.allclose(fr, HDR["fch1"] + HDR["foff"] * np.arange(10, 13), rtol=0, atol=1e-9)


def test_second_if_and_compressed_chunks(tmp_path):
    a = spec(777, 9, seed=2)
    FB.write_filterbank(tmp_path / "b.h5", np.stack([a, -a], axis=1), HDR, chunks=(4, 1, 128), deflate=6,
                        shuffle=True)
    with FB.FilterbankFile(tmp_path / "b.h5") as f:
        assert f.shape == (777, 2, 9)
        assert [i for i, _ in f.filters] == [2, 1]  # shuffle, deflate
        assert np.array_equal(f.read(ifnum=1), -a)
        assert np.array_equal(f.read(channels=range(5, 700), ifnum=0, times=range(2, 9)), a[5:700, 2:9])


def test_read_into_buffers(tmp_path):
    a = spec(300, 7, seed=3)
    FB.write_filterbank(tmp_path / "c.h5", a)
    out = np.zeros((120, 7), np.float32, order="F")
    assert FB.read_spectrogram(tmp_path / "c.h5", channels=range(50, 170), out=out) is out
    assert np.array_equal(out, a[50:170])
    # the layout pinned_spectrogram hands out: a transposed (Nt, Nf) tensor
    t = torch.empty((7, 120), dtype=torch.float32).t()
    FB.read_spectrogram(tmp_path / "c.h5", channels=range(50, 170), out=t)
    assert np.array_equal(t.numpy(), a[50:170])
    for bad in (np.zeros((120, 7), np.float32), np.zeros((121, 7), np.float32, order="F"),
                torch.zeros((120, 7)), torch.empty((7, 120), dtype=torch.float64).t()):
        with pytest.raises(ValueError):
            FB.read_spectrogram(tmp_path / "c.h5", channels=range(50, 170), out=bad)


def test_errors(tmp_path):
    a = spec(64, 4)
    # a filter libhdf5 cannot run here (blosc, registered as optional so the
    # write stores the chunks unfiltered): refused before any read
    FB.write_filterbank(tmp_path / "d.h5", a, extra_filter=(32001, True))
    with FB.FilterbankFile(tmp_path / "d.h5") as f:
        assert f.filters and f.filters[0][0] == 32001
        with pytest.raises(FB.FilterbankError, match="32001"):
            f.read()
    # bitshuffle alone is read by direct chunk reads without its plugin; these
    # chunks were stored unfiltered (the filter was skipped on write)
    FB.write_filterbank(tmp_path / "g.h5", a, extra_filter=(32008, True))
    with FB.FilterbankFile(tmp_path / "g.h5") as f:
        assert f.filters[0][0] == 32008 and np.array_equal(f.read(), a)
    FB.write_filterbank(tmp_path / "e.h5", a)
    with FB.FilterbankFile(tmp_path / "e.h5") as f:
        for kw in ({"channels": range(60, 70)}, {"channels": range(5, 5)}, {"times": range(0, 5)},
                   {"ifnum": 1}, {"channels": range(0, 10, 2)}):
            with pytest.raises(ValueError):
                f.read(**kw)
    with pytest.raises(FB.FilterbankError, match="no dataset"):
        FB.FilterbankFile(tmp_path / "e.h5", dataset="spectra")
    (tmp_path / "f.h5").write_bytes(b"not an hdf5 file")
    with pytest.raises(FB.FilterbankError, match="cannot open"):
        FB.FilterbankFile(tmp_path / "f.h5")
