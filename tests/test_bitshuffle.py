"""Bitshuffle chunk decoding (zdt_bshuf_decompress, csrc/bitshuffle.cpp) and
the filterbank ingest of bitshuffle-compressed HDF5 files without the filter
plugin (fdrt_hip.filterbank).  The reference reads such files through the
H5Zbitshuffle plugin (test/heavytests.jl:6).  CPU only: the decoder is host
code of the library.

What pins what:
  * the bit transposition and its blocking (block size, the partial last
    block, raw leftover elements) against vectors the bitshuffle library
    0.3.5 wrote (tests/golden/bitshuffle_0.3.5.npz, make_bitshuffle_vectors.py);
  * LZ4 block decoding against the LZ4 library (liblz4, when the system has
    it) compressing varied data;
  * the LZ4 container (12-byte big-endian header, size-prefixed blocks) is
    assembled here as the library's HDF5 filter writes it: that layout is
    not pinned by an independent encoder (no bitshuffle-LZ4 file or encoder
    is available here)."""
import ctypes as C
import ctypes.util
from pathlib import Path

import numpy as np
import pytest

from fdrt_hip import _lib
from fdrt_hip import filterbank as FB

GOLD = np.load(Path(__file__).parent / "golden" / "bitshuffle_0.3.5.npz")
NCASES = sum(1 for k in GOLD.files if k.endswith("_params"))


def decode(buf: bytes, out_len: int, E: int, block: int, comp: int):
    out = np.zeros(out_len, np.uint8)
    src = np.frombuffer(buf, np.uint8) if len(buf) else np.zeros(1, np.uint8)
    rc = _lib.lib().zdt_bshuf_decompress(src.ctypes.data_as(C.c_void_p), len(buf), out.ctypes.data_as(C.c_void_p),
                                         out_len, E, block, comp)
    return rc, out.tobytes()


def default_block(E):
    return max((8192 // E) // 8 * 8, 128)


def forward(raw: bytes, E: int, block: int) -> bytes:
    """Bit transposition as the library lays it out: per block of nb elements
    (nb a multiple of 8), plane 8b + k holds bit k of byte b of every element
    (element i at bit i mod 8 of byte i/8); leftover (< 8) elements raw."""
    a = np.frombuffer(raw, np.uint8).reshape(-1, E)
    n = a.shape[0]
    block = block or default_block(E)
    out, done = [], 0
    while n - done >= 8:
        nb = block if n - done >= block else (n - done) // 8 * 8
        bits = np.unpackbits(a[done:done + nb], axis=1, bitorder="little")  # (nb, 8E): byte b, bit k -> 8b + k
        out.append(np.packbits(bits.T, axis=1, bitorder="little").tobytes())
        done += nb
    out.append(a[done:].tobytes())
    return b"".join(out)


def blocks_of(n: int, E: int, block: int):
    block = block or default_block(E)
    sizes, done = [], 0
    while n - done >= 8:
        nb = block if n - done >= block else (n - done) // 8 * 8
        sizes.append(nb)
        done += nb
    return sizes, n - done


def _liblz4():
    for c in (ctypes.util.find_library("lz4"), "liblz4.so.1", "/opt/conda/lib/liblz4.so"):
        if not c:
            continue
        try:
            L = C.CDLL(c)
        except OSError:
            continue
        L.LZ4_compress_default.restype = C.c_int
        L.LZ4_compress_default.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.c_int]
        L.LZ4_compressBound.restype = C.c_int
        L.LZ4_compressBound.argtypes = [C.c_int]
        return L
    return None


LZ4 = _liblz4()


def lz4(data: bytes) -> bytes:
    cap = LZ4.LZ4_compressBound(len(data))
    dst = C.create_string_buffer(cap)
    k = LZ4.LZ4_compress_default(data, dst, len(data), cap)
    assert k > 0
    return dst.raw[:k]


def lz4_literals(data: bytes) -> bytes:
    """A valid LZ4 block holding only literals (no compressor needed)."""
    n = len(data)
    if n < 15:
        return bytes([n << 4]) + data
    ext, r = [], n - 15
    while r >= 255:
        ext.append(255)
        r -= 255
    ext.append(r)
    return bytes([0xF0]) + bytes(ext) + data


def container(raw: bytes, E: int, block: int, compress=None) -> bytes:
    """The LZ4-mode chunk the HDF5 filter writes: uint64 BE size, uint32 BE
    block bytes, then per block uint32 BE compressed size + LZ4 block of its
    transposed bytes; the leftover elements raw."""
    compress = compress or (lz4 if LZ4 else lz4_literals)
    block = block or default_block(E)
    n = len(raw) // E
    sizes, rest = blocks_of(n, E, block)
    planes = forward(raw, E, block)
    out = [len(raw).to_bytes(8, "big"), (block * E).to_bytes(4, "big")]
    pos = 0
    for nb in sizes:
        c = compress(planes[pos:pos + nb * E])
        out += [len(c).to_bytes(4, "big"), c]
        pos += nb * E
    out.append(planes[pos:])
    return b"".join(out)


@pytest.mark.parametrize("q", range(NCASES))
def test_plain_bitshuffle_vs_library(q):
    E, n, bs = (int(x) for x in GOLD[f"case{q}_params"])
    raw, shuf = GOLD[f"case{q}_raw"].tobytes(), GOLD[f"case{q}_shuffled"].tobytes()
    assert forward(raw, E, bs) == shuf  # the test's own encoder is the library's layout
    rc, out = decode(shuf, len(raw), E, bs, 0)
    assert rc == 0 and out == raw


@pytest.mark.parametrize("q", range(NCASES))
def test_lz4_container(q):
    E, n, bs = (int(x) for x in GOLD[f"case{q}_params"])
    raw = GOLD[f"case{q}_raw"].tobytes()
    rc, out = decode(container(raw, E, bs), len(raw), E, 0, 2)
    assert rc == 0 and out == raw
    rc, out = decode(container(raw, E, bs, lz4_literals), len(raw), E, 0, 2)
    assert rc == 0 and out == raw


@pytest.mark.skipif(LZ4 is None, reason="liblz4 not on this system")
def test_lz4_decoder_vs_liblz4():
    """Compressible inputs: long zero runs (long matches, overlapping copies,
    255-continued lengths), repeats at far offsets, short periods, mixed with
    random literals, all through the library's own compressor."""
    rng = np.random.default_rng(7)
    parts = [bytes(70000), rng.integers(0, 256, 3000, dtype=np.uint8).tobytes(), b"abc" * 5000,
             np.repeat(rng.integers(0, 256, 500, dtype=np.uint8), 37).tobytes()]
    far = rng.integers(0, 256, 40000, dtype=np.uint8).tobytes()
    parts.append(far + far[:20000])  # a match 40000 bytes back
    for data in parts + [b"".join(parts)]:
        data = data[: len(data) // 8 * 8]
        rc, out = decode(container(data, 1, len(data)), len(data), 1, 0, 2)
        assert rc == 0 and out == data


def test_malformed_chunks_are_refused():
    raw = GOLD["case0_raw"].tobytes()
    good = container(raw, 4, 0)
    assert decode(good, len(raw), 4, 0, 2)[0] == 0
    bad = [good[:11], good[:-1], good + b"\0", b"\0" * 8 + good[8:],  # truncated, trailing, wrong size
           good[:12] + (10 ** 6).to_bytes(4, "big") + good[16:],   # a block past the end
           good[:8] + (3).to_bytes(4, "big") + good[12:]]          # block bytes not whole elements
    for b in bad:
        assert decode(b, len(raw), 4, 0, 2)[0] == _lib.ZDT_EINVAL
    # an LZ4 match offset of 0, and one before the start of the output
    for blk in (bytes([0x14]) + b"x" + b"\0\0", bytes([0x14]) + b"x" + b"\x05\x00"):
        c = (8).to_bytes(8, "big") + (8).to_bytes(4, "big") + len(blk).to_bytes(4, "big") + blk
        assert decode(c, 8, 1, 0, 2)[0] == _lib.ZDT_EINVAL
    assert decode(good, len(raw), 4, 0, 3)[0] == _lib.ZDT_ENOTSUP  # zstd


# ------------------------------------------------------------------ HDF5 files


def write_bitshuffled(path, data, chunks, comp=2, block=0, skip=()):
    """A filterbank file whose chunks carry filter 32008 (cd = 0, 3, 4, block,
    comp), written with direct chunk writes as the bitshuffle filter would;
    chunk offsets in ``skip`` are stored unfiltered (filter mask bit set)."""
    h = FB._lib()
    a = np.asarray(data, np.float32)
    if a.ndim == 2:
        a = a[:, None, :]
    nchans, nifs, ntime = a.shape
    tic = np.ascontiguousarray(np.transpose(a, (2, 1, 0)))  # (time, if, chan)
    f = h.L.H5Fcreate(str(path).encode(), 2, 0, 0)
    sp = h.ok(h.L.H5Screate_simple(3, (FB._hsize * 3)(ntime, nifs, nchans), None), "H5Screate_simple")
    dcpl = h.ok(h.L.H5Pcreate(h.DCPL), "H5Pcreate")
    h.ok(h.L.H5Pset_chunk(dcpl, 3, (FB._hsize * 3)(*chunks)), "H5Pset_chunk")
    cd = (C.c_uint * 5)(0, 3, 4, block, comp)
    h.ok(h.L.H5Pset_filter(dcpl, 32008, 1, 5, cd), "H5Pset_filter")  # optional: libhdf5 has no plugin
    d = h.ok(h.L.H5Dcreate2(f, b"data", h.IEEE_F32LE, sp, 0, dcpl, 0), "H5Dcreate2")
    ct, ci, cc = chunks
    for t in range(0, ntime, ct):
        for i in range(0, nifs, ci):
            for c in range(0, nchans, cc):
                blk = np.zeros(chunks, np.float32)  # edge chunks are stored whole
                sub = tic[t:t + ct, i:i + ci, c:c + cc]
                blk[:sub.shape[0], :sub.shape[1], :sub.shape[2]] = sub
                rawb = blk.tobytes()
                if (t, i, c) in skip:
                    payload, mask = rawb, 1
                elif comp == 2:
                    payload, mask = container(rawb, 4, block), 0
                else:
                    payload, mask = forward(rawb, 4, block), 0
                buf = np.frombuffer(payload, np.uint8)
                h.ok(h.L.H5Dwrite_chunk(d, 0, mask, (FB._hsize * 3)(t, i, c), len(payload),
                                        buf.ctypes.data_as(C.c_void_p)), "H5Dwrite_chunk")
    FB._write_attr(h, d, "nchans", nchans)
    FB._write_attr(h, d, "foff", -1.0)
    FB._write_attr(h, d, "fch1", 1000.0)
    h.L.H5Dclose(d)
    h.L.H5Pclose(dcpl)
    h.L.H5Sclose(sp)
    h.L.H5Fclose(f)


@pytest.mark.parametrize("comp,block", [(2, 0), (2, 512), (0, 0)])
def test_filterbank_bitshuffle_without_plugin(tmp_path, comp, block):
    rng = np.random.default_rng(comp * 1000 + block)
    a = rng.exponential(1.0, size=(3000, 2, 21)).astype(np.float32)  # Nf x nifs x Nt
    write_bitshuffled(tmp_path / "b.h5", a, (8, 1, 1024), comp, block, skip={(8, 1, 2048)})
    with FB.FilterbankFile(tmp_path / "b.h5") as f:
        assert f.shape == (3000, 2, 21) and [i for i, _ in f.filters] == [32008]
        assert f._bshuf is not None
        assert np.array_equal(f.read(), a[:, 0, :])
        assert np.array_equal(f.read(ifnum=1), a[:, 1, :])
        # hyperslabs across chunk edges, the unfiltered chunk included
        assert np.array_equal(f.read(channels=range(1000, 2900), ifnum=1, times=range(5, 19)), a[1000:2900, 1, 5:19])
        assert np.array_equal(f.read(channels=range(2047, 2049), times=range(20, 21)), a[2047:2049, 0, 20:21])
        out = np.zeros((3000, 21), np.float32, order="F")
        assert FB.read_spectrogram(tmp_path / "b.h5", out=out) is out and np.array_equal(out, a[:, 0, :])


def test_bitshuffle_with_other_filters_is_refused(tmp_path):
    # bitshuffle after deflate: not decodable here without the plugin
    FB.write_filterbank(tmp_path / "c.h5", np.ones((64, 4), np.float32), deflate=1, extra_filter=(32008, True))
    with FB.FilterbankFile(tmp_path / "c.h5") as f:
        assert f._bshuf is None
        with pytest.raises(FB.FilterbankError, match="32008"):
            f.read()
