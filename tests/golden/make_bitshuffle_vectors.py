"""Writes tests/golden/bitshuffle_0.3.5.npz: inputs and their bit-transposed
form as the bitshuffle library (0.3.5, through imagecodecs 2021.8.26's
bitshuffle_encode, which calls bshuf_bitshuffle) produces it, for the
decoder zdt_bshuf_decompress (csrc/bitshuffle.cpp).

The cases cover element sizes 1, 2, 4 (Float32, the filterbank data) and 8,
the default block size (8 KiB of elements) and explicit ones, a partial last
block (cut to a multiple of 8 elements) and fewer than 8 leftover elements
(stored raw).

Run once, in this container, with the interpreter that has imagecodecs:
    /opt/conda/bin/python3.9 tests/golden/make_bitshuffle_vectors.py
"""
from pathlib import Path

import imagecodecs
import numpy as np

CASES = [  # (elem_size, n elements, block size in elements; 0 = the library's default)
    (4, 2048, 0),
    (4, 5000 + 5, 0),
    (4, 1003, 256),
    (4, 1000, 256),
    (4, 7, 0),
    (2, 13, 0),
    (2, 4099, 64),
    (1, 777, 128),
    (8, 600, 0),
]


def main():
    rng = np.random.default_rng(32008)
    out = {"version": np.array(imagecodecs.bitshuffle_version())}
    for q, (E, n, bs) in enumerate(CASES):
        if E == 4:  # float32 powers, like a spectrogram's
            raw = rng.exponential(1.0, n).astype(np.float32).tobytes()
        else:
            raw = rng.integers(0, 256, n * E, dtype=np.uint8).tobytes()
        enc = imagecodecs.bitshuffle_encode(raw, itemsize=E, blocksize=bs)
        assert len(enc) == len(raw)
        assert bytes(imagecodecs.bitshuffle_decode(enc, itemsize=E, blocksize=bs)) == raw
        out[f"case{q}_params"] = np.array([E, n, bs], np.int64)
        out[f"case{q}_raw"] = np.frombuffer(raw, np.uint8)
        out[f"case{q}_shuffled"] = np.frombuffer(bytes(enc), np.uint8)
    path = Path(__file__).parent / "bitshuffle_0.3.5.npz"
    np.savez_compressed(path, **out)
    print(path, path.stat().st_size, "bytes")


if __name__ == "__main__":
    main()
