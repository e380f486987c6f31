"""Seeded synthetic spectrograms (test infrastructure only).

A spectrogram is Float32 (Nf, Nt), frequency fastest: i.i.d. Exponential(1)
noise plus linearly drifting tones, each adding amplitude ``a`` at channel
round(c0 + r*t) mod Nf for every time step t.
"""
from __future__ import annotations

from typing import Sequence, Tuple

import numpy as np


def spectrogram(Nf: int, Nt: int, tones: Sequence[Tuple[float, float, float]],
                seed: int = 1234, noise: bool = True) -> np.ndarray:
    """Fortran-ordered float32 (Nf, Nt).  ``tones`` holds (c0, rate, amp)."""
    rng = np.random.default_rng(seed)
    if noise:
        s = rng.exponential(1.0, size=(Nt, Nf)).astype(np.float32).T
    else:
        s = np.zeros((Nt, Nf), dtype=np.float32).T
    s = np.asfortranarray(s)
    for c0, r, a in tones:
        for t in range(Nt):
            s[int(round(c0 + r * t)) % Nf, t] += np.float32(a)
    return s


def random_tones(Nf: int, n: int, amp: Tuple[float, float], rate: Tuple[float, float],
                 seed: int) -> list:
    rng = np.random.default_rng(seed + 7919)
    c0 = rng.uniform(0, Nf, n)
    r = rng.uniform(rate[0], rate[1], n)
    a = rng.uniform(amp[0], amp[1], n)
    return list(zip(c0.tolist(), r.tolist(), a.tolist()))


def cfg1(seed: int = 1234) -> np.ndarray:
    """2^16 x 16, three tones of amplitude 20 at rates +0.5, -1.25, +3.0."""
    Nf, Nt = 65536, 16
    rng = np.random.default_rng(seed + 1)
    c0 = rng.uniform(0, Nf, 3)
    tones = [(c0[0], 0.5, 20.0), (c0[1], -1.25, 20.0), (c0[2], 3.0, 20.0)]
    return spectrogram(Nf, Nt, tones, seed)


def cfg2(seed: int = 1234, Nf: int = 1 << 20, Nt: int = 16) -> np.ndarray:
    """2^20 x 16, 16 tones, amplitude U(10,50), rate U(-4,4)."""
    return spectrogram(Nf, Nt, random_tones(Nf, 16, (10, 50), (-4, 4), seed), seed)
