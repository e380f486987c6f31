"""Host-side arithmetic of the ZDT path, restated for checking (test
infrastructure only; nothing in the product imports it).

Functions and what pins them (tests/test_oracle.py):

``calcNl``/``growNr``  nextprod((2,3,5), Nt+Nr-1) (src/zdtutils.jl:7-19).
                      Pinned by the 3x2 case (Nl=4) and the sizes quoted for
                      the reference's configurations (144, 48, 288, 540, 1125)
                      and a brute-force factor check.
``estimate_memory``   4*(Nf*Nt*(Ni+3) + Nf*Nl*3 + Nf*Nr*No) (src/zdtutils.jl:35-53).
                      Pinned by hand-evaluated values.
``batchrates``        src/batchrates.jl:22-60 with exact rationals.  Pinned by
                      hand-derived cases and its own invariants (equal lengths,
                      divisibility, coverage of the requested span).
``fdrstats``          mean of column 1, minimum over columns of the corrected
                      standard deviation (src/fdrutils.jl:39-50).  Pinned
                      against numpy.mean/numpy.std(ddof=1) (textbook routines).
``fdrdenormalize``/``fdrnormalize``  snr*s+m and (x-m)/s (src/fdrutils.jl:65-119).
``findprotohits``     findall(>=(thr)) in column-major order, 1-based, with the
                      multi-batch column offset (src/fdrutils.jl:148-169).
                      Pinned against numpy.nonzero on the transposed matrix.
"""
from __future__ import annotations

import math
from fractions import Fraction
from typing import List, Sequence, Tuple, Union

import numpy as np


def _smooth(n: int, factors) -> bool:
    for f in factors:
        while n % f == 0 and n > 1:
            n //= f
    return n == 1


def nextprod(factors, n: int) -> int:
    """Smallest integer >= n whose prime factors all lie in ``factors``."""
    n = max(int(n), 1)
    m = n
    while not _smooth(m, factors):
        m += 1
    return m


def calcNl(Nt: int, Nr: int, factors=(2, 3, 5)) -> int:
    """src/zdtutils.jl:7-9"""
    return nextprod(factors, Nt + Nr - 1)


def growNr(Nt: int, Nr: int, factors=(2, 3, 5)) -> int:
    """src/zdtutils.jl:17-19"""
    return calcNl(Nt, Nr, factors) - Nt + 1


def estimate_memory(Nf: int, Nt: int, Nr: int, Ni: int = 1, No: int = 1,
                    factors=(2, 3, 5)) -> int:
    """src/zdtutils.jl:35-53"""
    Nl = calcNl(Nt, Nr, factors)
    Ni += 3
    Nz = 3
    return 4 * ((Nf * Nt) * Ni + (Nf * Nl) * Nz + (Nf * Nr) * No)


def _round_from_zero(x: float) -> int:
    return int(math.copysign(math.ceil(abs(x)), x)) if x != int(x) else int(x)


def _round_to_zero(x: float) -> int:
    return int(math.trunc(x))


def batchrates(Nt: int, dhzps, r1, r2=None, Nrb: int = 360
               ) -> List[Tuple[Fraction, Fraction, int]]:
    """src/batchrates.jl:22-60.  Returns (first, step, length) triples of
    exact rationals, one per batch."""
    if r2 is None:
        r2 = -r1
    if r2 < r1:
        r1, r2 = r2, r1
    nc1 = int(round(r1 / dhzps))          # round(Int, x): ties to even
    nc2 = int(round(r2 / dhzps))
    Nrspan = abs(nc2 - nc1) + 1
    Nrb = min(Nrb, Nrspan)
    Nrb = growNr(Nt, Nrb)
    Nr = -(-Nrspan // Nrb) * Nrb             # cld(Nrspan, Nrb) * Nrb
    Nrbonus = Nr - Nrspan
    sgn = float(np.sign(dhzps))
    nc1 -= _round_from_zero(sgn * Nrbonus / 2)
    nc2 += _round_to_zero(sgn * Nrbonus / 2)
    first = Fraction(nc1, Nt - 1)
    last = Fraction(nc2, Nt - 1)
    step = (last - first) / (Nr - 1) if Nr > 1 else Fraction(0)
    assert Nr % Nrb == 0
    return [(first + b * Nrb * step, step, Nrb) for b in range(Nr // Nrb)]


# ---------------------------------------------------------------------------
# statistics and proto-hits
# ---------------------------------------------------------------------------


def _as_list(fdrs) -> List[np.ndarray]:
    if isinstance(fdrs, np.ndarray) and fdrs.ndim == 2:
        return [fdrs]
    return [np.asarray(f) for f in fdrs]


def fdrstats(fdrs) -> Tuple[float, float]:
    """(mean of column 1 of the first matrix, min over all columns of the
    Bessel-corrected std), float64 (src/fdrutils.jl:39-50)."""
    lst = _as_list(fdrs)
    m = float(np.mean(lst[0][:, 0].astype(np.float64)))
    s = min(float(np.min(np.std(f.astype(np.float64), axis=0, ddof=1))) for f in lst)
    return m, s


def fdrdenormalize(snr, m, s) -> float:
    """snr*s + m (src/fdrutils.jl:104-107)."""
    return float(snr) * float(s) + float(m)


def fdrnormalize(fdrs, m=None, s=None):
    """(x - m)/s for one matrix or a list (src/fdrutils.jl:65-90)."""
    lst = _as_list(fdrs)
    if m is None:
        m, s = fdrstats(lst)
    out = [((f.astype(np.float64) - m) / s).astype(np.float32) for f in lst]
    return out[0] if isinstance(fdrs, np.ndarray) and fdrs.ndim == 2 else out


def findprotohits(fdrs, threshold, snr: bool = False) -> np.ndarray:
    """1-based (i, j) of every element >= threshold, column-major order;
    batches are treated as hcat'd (src/fdrutils.jl:148-169)."""
    lst = _as_list(fdrs)
    if snr:
        m, s = fdrstats(lst)
        threshold = fdrdenormalize(threshold, m, s)
    Nrb = lst[0].shape[1]
    out = []
    for b, f in enumerate(lst):
        jj, ii = np.nonzero(f.astype(np.float64).T >= threshold)
        out.append(np.stack([ii + 1, jj + 1 + b * Nrb], axis=1))
    if not out:
        return np.zeros((0, 2), dtype=np.int64)
    return np.concatenate(out).astype(np.int64)
