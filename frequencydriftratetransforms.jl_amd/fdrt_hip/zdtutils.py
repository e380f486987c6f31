"""Sizing and rate batching (host side of the boundary).

calcNl/growNr/estimate_memory call the C-ABI (src/zdtutils.jl:7-53);
batchrates restates src/batchrates.jl:22-60 with exact rationals, returning
``RateRange`` objects that stand in for Julia's rational ranges.
"""
from __future__ import annotations

import math
from fractions import Fraction
from typing import List, Sequence, Union

import numpy as np

from . import _lib


def _factors(factors):
    if factors is None:
        factors = (2, 3, 5)
    return _lib.int_array(factors)


def calcNl(Nt: int, Nr: int, factors=(2, 3, 5)) -> int:
    """nextprod(factors, Nt + Nr - 1)  (src/zdtutils.jl:7-9)."""
    arr, n = _factors(factors)
    v = _lib.lib().zdt_calc_nl(int(Nt), int(Nr), arr, n)
    if v < 0:
        raise ValueError("calcNl: invalid arguments")
    return int(v)


def growNr(Nt: int, Nr: int, factors=(2, 3, 5)) -> int:
    """calcNl(Nt, Nr, factors) - Nt + 1  (src/zdtutils.jl:17-19)."""
    arr, n = _factors(factors)
    v = _lib.lib().zdt_grow_nr(int(Nt), int(Nr), arr, n)
    if v < 0:
        raise ValueError("growNr: invalid arguments")
    return int(v)


def estimate_memory(Nf: int, Nt: int, Nr: int, Ni: int = 1, No: int = 1, factors=(2, 3, 5)) -> int:
    """Bytes per the reference's formula (src/zdtutils.jl:35-53)."""
    arr, n = _factors(factors)
    return int(_lib.lib().zdt_estimate_memory(int(Nf), int(Nt), int(Nr), int(Ni), int(No), arr, n))


def workspace_bytes(Nf: int, Nt: int, Nr: int, factors=(2, 3, 5), generic: bool = False) -> int:
    """Device bytes this library allocates for such a workspace."""
    arr, n = _factors(factors)
    return int(_lib.lib().zdt_workspace_bytes(int(Nf), int(Nt), int(Nr), arr, n,
                                              _lib.ZDT_FLAG_GENERIC if generic else 0))


class RateRange:
    """An evenly spaced range of normalized drift rates (channels per time
    step), the counterpart of a Julia ``StepRangeLen``: first, step, length
    held exactly."""

    def __init__(self, first, step, length: int):
        self.first = Fraction(first)
        self.step = Fraction(step)
        self.length = int(length)

    def __len__(self) -> int:
        return self.length

    def __getitem__(self, i):
        if isinstance(i, slice):
            return [self[k] for k in range(*i.indices(self.length))]
        if i < 0:
            i += self.length
        if not 0 <= i < self.length:
            raise IndexError(i)
        return self.first + i * self.step

    def __iter__(self):
        for i in range(self.length):
            yield self.first + i * self.step

    def __eq__(self, other) -> bool:
        return list(self) == list(other)

    def __repr__(self) -> str:
        return f"RateRange({self.first}, step={self.step}, length={self.length})"

    def tolist(self) -> List[float]:
        return [float(x) for x in self]


def driftrates(ws, r0=None) -> RateRange:
    """range(r0, step=ws.dr, length=ws.Nr)  (src/zdtutils.jl:62-64)."""
    r0 = ws.r0 if r0 is None else np.float32(r0)
    return RateRange(Fraction(float(r0)), Fraction(float(ws.dr)), ws.Nr)


def _round_from_zero(x: float) -> int:
    return int(math.copysign(math.ceil(abs(x)), x))


def batchrates(Nt: int, dhzps, r1, r2=None, Nrb: int = 360) -> List[RateRange]:
    """Equal-length batches of normalized drift rates spanning r1..r2 Hz/s in
    steps of dhzps Hz/s for Nt time steps (src/batchrates.jl:22-60)."""
    if r2 is None:
        r2 = -r1
    if r2 < r1:
        r1, r2 = r2, r1
    nc1 = int(round(r1 / dhzps))  # round(Int, x): nearest, ties to even
    nc2 = int(round(r2 / dhzps))
    Nrspan = abs(nc2 - nc1) + 1
    Nrb = min(Nrb, Nrspan)
    Nrb = growNr(Nt, Nrb)
    Nr = -(-Nrspan // Nrb) * Nrb  # cld(Nrspan, Nrb) * Nrb
    Nrbonus = Nr - Nrspan
    half = float(np.sign(dhzps)) * Nrbonus / 2
    nc1 -= _round_from_zero(half)
    nc2 += int(math.trunc(half))
    first = Fraction(nc1, Nt - 1)
    last = Fraction(nc2, Nt - 1)
    step = (last - first) / (Nr - 1) if Nr > 1 else Fraction(0)
    if Nr % Nrb != 0:
        raise AssertionError(f"number of rates ({Nr}) is not divisible by Nrb ({Nrb})")
    return [RateRange(first + b * Nrb * step, step, Nrb) for b in range(Nr // Nrb)]


def float32_of(x) -> np.float32:
    """Julia's Float32(x): Rational -> Float32(num)/Float32(den)."""
    if isinstance(x, Fraction):
        return np.float32(np.float32(x.numerator) / np.float32(x.denominator))
    return np.float32(x)


def rates_params(rates) -> tuple:
    """(r0, dr, Nr) of an evenly spaced rate collection (first/step/length)."""
    if isinstance(rates, RateRange):
        return float32_of(rates.first), float32_of(rates.step), rates.length
    if isinstance(rates, range):
        return np.float32(rates.start), np.float32(rates.step), len(rates)
    vals = list(rates)
    if not vals:
        raise ValueError("rates must not be empty")
    if len(vals) == 1:
        return float32_of(vals[0]), np.float32(0), 1
    step = vals[1] - vals[0]
    for i in range(2, len(vals)):
        d = vals[i] - vals[i - 1]
        if isinstance(d, Fraction) or isinstance(step, Fraction):
            ok = d == step
        else:
            ok = math.isclose(float(d), float(step), rel_tol=1e-6, abs_tol=1e-12)
        if not ok:
            raise ValueError("the ZDT requires evenly spaced drift rates")
    return float32_of(vals[0]), float32_of(step), len(vals)
