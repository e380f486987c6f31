"""CPU oracle for the ZDT transform (test infrastructure only).

Nothing in the product imports this module: only the tests, the smoke check
and the CPU-baseline leg of the benchmark use it, as the checker.

Notation follows the reference (zero-based here): the spectrogram ``s`` has
shape (Nf, Nt), frequency fastest; ``Nk = Nf//2 + 1``; rates are
``r_j = r0 + j*dr`` for ``j < Nr`` with ``r0``/``dr`` stored as Float32
(src/zdtfdr.jl:13-14).  Every FDR returned here is *normalized* (the inverse
transforms carry their 1/N), i.e. the drift-line sum; the reference's source
reads that way (plan_irfft/plan_ifft! are scaled plans, src/zdtfdr.jl:120,123,
and the CUDA extension multiplies by ``irfft_plan.scale``,
ext/CUDAFrequencyDriftRateTransformsExt.jl:101).  The reference's tests divide
by Nf*Nl (test/runtests.jl:31) which is the unnormalized convention; that
factor is available as ``unnormalized_scale``.

Functions and what pins them (tests/test_oracle.py):

``fdr64``      FP64 definition, FDR[:,j] = irfft(sum_t F[:,t] e^{2 pi i k t r_j/Nf})
               with phases reduced exactly in integer arithmetic
               (src/fftfdr.jl:10-12,86-96 restated for a rate range).
               Pinned by: the 3x2 known answer (test/runtests.jl:5-16); the
               integer-rate closed form ``intfdr``; brute-force O(Nf^2) DFT
               sums on tiny inputs; time-reversal and linearity invariants.
``zdt64``      FP64 restatement of the reference's Bluestein chain
               (src/zdtfdr.jl:140-331) with exact phases.  Pinned by equality
               with ``fdr64`` (independent algorithm), independence of Nl
               (factors), and the 3x2 known answer.
``zdt_ref32``  FP32 restatement of src/zdtfdr.jl following Julia's Float32
               promotion order for every phase argument
               (src/zdtfdr.jl:141,197,269) and complex64 FFTs.  Pinned by the
               3x2 known answer and by its distance to ``fdr64`` (bounded by
               the Float32 phase-argument rounding it reproduces).
``intshift``/``intfdr``  brute force (src/intfdr.jl:7-14,34-45).  Pinned by
               the known answers of test/runtests.jl:8-20.
``fdshift``    FP64 Fourier-domain shift of every time step
               (src/fftfdr.jl:104-118).  Pinned by test/runtests.jl:25 and by
               integer-shift equality with ``intshift``.
"""
from __future__ import annotations

import math
from fractions import Fraction
from typing import Callable, Iterable, Optional, Sequence, Union

import numpy as np
import scipy.fft as sfft

Window = Union[None, str, Callable[[int, int], float], np.ndarray]

# ---------------------------------------------------------------------------
# exact phase arithmetic
# ---------------------------------------------------------------------------


def f32(x) -> np.float32:
    """Round to Float32 the way Julia's ``Float32(x)`` does (nearest-even)."""
    if isinstance(x, Fraction):
        # Julia: Float32(num)/Float32(den) (base/rational.jl), exact operands
        return np.float32(np.float32(x.numerator) / np.float32(x.denominator))
    return np.float32(x)


def exact(x) -> Fraction:
    """Exact rational value of a binary float (or a Fraction)."""
    if isinstance(x, Fraction):
        return x
    return Fraction(float(x))


def _int_array(a) -> np.ndarray:
    return np.asarray(a, dtype=np.int64)


def _mul_exact(a: np.ndarray, b) -> np.ndarray:
    """a*b exactly: int64 when it cannot overflow, Python ints otherwise."""
    amax = int(np.max(np.abs(a))) if np.size(a) else 0
    bmax = int(np.max(np.abs(np.asarray(b, dtype=object)))) if np.size(b) else 0
    if amax * bmax < (1 << 62) and not (isinstance(a, np.ndarray) and a.dtype == object):
        return a * np.asarray(b, dtype=np.int64)
    return np.asarray(a, dtype=object) * np.asarray(b, dtype=object)


def _revs(num: np.ndarray, den: int) -> np.ndarray:
    """(num mod den)/den wrapped to [-0.5, 0.5), in float64."""
    m = np.mod(num, den)
    r = np.asarray(m, dtype=np.float64) / float(den)
    r = np.where(r >= 0.5, r - 1.0, r)
    return r


def cis_revs(rev: np.ndarray) -> np.ndarray:
    """exp(2*pi*i*rev) in complex128, with rev already reduced."""
    ang = 2.0 * np.pi * rev
    return np.cos(ang) + 1j * np.sin(ang)


def _common_den(*xs: Fraction) -> int:
    q = 1
    for x in xs:
        q = q * x.denominator // math.gcd(q, x.denominator)
    return q


# ---------------------------------------------------------------------------
# windows
# ---------------------------------------------------------------------------


def window64(w: Window, Nf: int) -> Optional[np.ndarray]:
    """Window values w(k, Nf) for k < Nk in float64; None for rectangular.

    ``"hamming"`` is 0.53836 + 0.46164*cospi(2k/Nf) (src/zdtfdr.jl:303-305);
    a callable is evaluated as the reference does (src/zdtfdr.jl:297-301).
    """
    Nk = Nf // 2 + 1
    if w is None or (isinstance(w, str) and w == "rect"):
        return None
    if isinstance(w, str):
        if w == "hamming":
            k = np.arange(Nk, dtype=np.float64)
            return 0.53836 + 0.46164 * np.cos(np.pi * (2.0 * k / Nf))
        raise ValueError(f"unsupported window type ({w})")
    if callable(w):
        return np.array([float(w(k, Nf)) for k in range(Nk)], dtype=np.float64)
    arr = np.asarray(w, dtype=np.float64)
    if arr.shape != (Nk,):
        raise ValueError("window vector must have Nf//2+1 entries")
    return arr


# ---------------------------------------------------------------------------
# FP64 definition oracle
# ---------------------------------------------------------------------------


def fdr64(spec: np.ndarray, r0, dr=None, Nr: Optional[int] = None,
          window: Window = None, rates: Optional[Sequence] = None) -> np.ndarray:
    """FP64 FDR by the Fourier-shift definition with exactly reduced phases.

    FDR[:, j] = irfft(w(k) * sum_t F[k, t] * exp(2*pi*i*k*t*r_j/Nf)), F = rfft(s).
    Restates src/fftfdr.jl:86-96 (phasor multiply, sum over time, irfft) for
    the rates r_j = r0 + j*dr with r0, dr taken as their Float32 values
    (src/zdtfdr.jl:13-14, 81-84), or for an explicit ``rates`` list of exact
    values.  Returns float64 (Nf, Nr).
    """
    spec = np.asarray(spec)
    Nf, Nt = spec.shape
    Nk = Nf // 2 + 1
    if rates is None:
        R0, D = exact(f32(r0)), exact(f32(dr))
        rl = [R0 + j * D for j in range(Nr)]
    else:
        rl = [exact(r) for r in rates]
    F = np.fft.rfft(spec.astype(np.float64), axis=0)
    k = np.arange(Nk, dtype=np.int64)
    w = window64(window, Nf)
    out = np.empty((Nf, len(rl)), dtype=np.float64)
    for j, r in enumerate(rl):
        den = Nf * r.denominator
        S = F[:, 0].copy()  # t = 0: phase 1
        if Nt > 1 and den < (1 << 61):
            # the phase numerator of (k, t) is t*step_k mod den, stepped
            # exactly in int64 (every value stays below 2*den < 2^62)
            step = np.mod(_mul_exact(k, r.numerator), den).astype(np.int64)
            acc = np.zeros(Nk, dtype=np.int64)
            for t in range(1, Nt):
                acc += step
                acc[acc >= den] -= den
                S += F[:, t] * cis_revs(_revs(acc, den))
        elif Nt > 1:
            t = np.arange(1, Nt, dtype=np.int64)
            num = _mul_exact(k[:, None] * t[None, :], r.numerator)
            S += np.sum(F[:, 1:] * cis_revs(_revs(num, den)), axis=1)
        if w is not None:
            S = S * w
        out[:, j] = np.fft.irfft(S, n=Nf)
    return out


def fftfdr64(spec: np.ndarray, rates: Iterable) -> np.ndarray:
    """FP64 ``fftfdr`` (src/fftfdr.jl:126-148): each rate rounded to Float32
    first, as ``phasor`` receives ``Float32(rate)`` (src/fftfdr.jl:10-12,88-89)."""
    return fdr64(spec, None, rates=[exact(f32(r)) for r in rates])


def fdshift64(spec: np.ndarray, rate) -> np.ndarray:
    """FP64 de-doppler of every time step by ``rate*t`` channels
    (``fdshift``, src/fftfdr.jl:104-118), exact phases."""
    spec = np.asarray(spec)
    Nf, Nt = spec.shape
    Nk = Nf // 2 + 1
    r = exact(f32(rate))
    F = np.fft.rfft(spec.astype(np.float64), axis=0)
    k = np.arange(Nk, dtype=np.int64)
    t = np.arange(Nt, dtype=np.int64)
    num = _mul_exact(k[:, None] * t[None, :], r.numerator)
    ph = cis_revs(_revs(num, Nf * r.denominator))
    return np.fft.irfft(F * ph, n=Nf, axis=0)


# ---------------------------------------------------------------------------
# FP64 Bluestein restatement
# ---------------------------------------------------------------------------


def _calcNl(Nt: int, Nr: int, factors=(2, 3, 5)) -> int:
    from .host_oracle import calcNl
    return calcNl(Nt, Nr, factors)


def zdt64(spec: np.ndarray, r0, dr, Nr: int, factors=(2, 3, 5),
          window: Window = None) -> np.ndarray:
    """FP64 restatement of the reference's ZDT chain with exact phases.

    input!      F = rfft(s)                                   src/zdtfdr.jl:209-212
    preprocess! Y[:, :Nt] = F * cispi(k t (t dr + 2 r0)/Nf), rest 0   :196-235
    computeV!   V = fft_l(vlow | vhigh)                       :140-187
    convolve!   Y = ifft_l(fft_l(Y) * V)                      :249-253
    postprocess! Ys = Y[:, :Nr] * cispi(k j^2 dr/Nf) * w(k)   :268-321
    output!     FDR = irfft(Ys, Nf)                           :328-331
    """
    spec = np.asarray(spec)
    Nf, Nt = spec.shape
    Nk = Nf // 2 + 1
    Nl = _calcNl(Nt, Nr, factors)
    R0, D = exact(f32(r0)), exact(f32(dr))
    Q = _common_den(R0, D)
    P0, PD = int(R0 * Q), int(D * Q)
    k = np.arange(Nk, dtype=np.int64)[:, None]
    F = np.fft.rfft(spec.astype(np.float64), axis=0)
    # preprocess: exponent i*pi*k*t*(t*dr + 2*r0)/Nf  -> revs k*t*(t*PD+2*P0)/(2*Nf*Q)
    t = np.arange(Nt, dtype=np.int64)[None, :]
    pre = cis_revs(_revs(_mul_exact(k * t, t * PD + 2 * P0), 2 * Nf * Q))
    Y = np.zeros((Nk, Nl), dtype=np.complex128)
    Y[:, :Nt] = F * pre
    # V: chirp exp(-i*pi*k*l^2*dr/Nf) at lags 0..Nr-1 and -(Nt-1)..-1
    V0 = np.zeros((Nk, Nl), dtype=np.complex128)
    lo = np.arange(Nr, dtype=np.int64)[None, :]
    V0[:, :Nr] = cis_revs(_revs(-_mul_exact(k * lo * lo, PD), 2 * Nf * Q))
    if Nt > 1:
        hi_cols = np.arange(Nl - Nt + 1, Nl, dtype=np.int64)
        lag = (Nl - hi_cols)[None, :]
        V0[:, Nl - Nt + 1:] = cis_revs(_revs(-_mul_exact(k * lag * lag, PD), 2 * Nf * Q))
    V = np.fft.fft(V0, axis=1)
    Z = np.fft.ifft(np.fft.fft(Y, axis=1) * V, axis=1)
    j = np.arange(Nr, dtype=np.int64)[None, :]
    post = cis_revs(_revs(_mul_exact(k * j * j, PD), 2 * Nf * Q))
    Ys = Z[:, :Nr] * post
    w = window64(window, Nf)
    if w is not None:
        Ys = Ys * w[:, None]
    return np.fft.irfft(Ys, n=Nf, axis=0)


# ---------------------------------------------------------------------------
# FP32 restatement of the reference (Julia Float32 promotion order)
# ---------------------------------------------------------------------------


def cispi32(arg: np.ndarray) -> np.ndarray:
    """Julia ``cispi(::Float32)``: exact reduction of the Float32 argument,
    then cos/sin rounded to Float32 (to within an ulp of Julia's kernels)."""
    a = np.asarray(arg, dtype=np.float32).astype(np.float64)
    n = np.rint(2.0 * a)
    f = a - 0.5 * n                      # exact, |f| <= 1/4
    s0 = np.sin(np.pi * f)
    c0 = np.cos(np.pi * f)
    q = np.mod(n, 4.0)
    c = np.select([q == 0, q == 1, q == 2], [c0, -s0, -c0], s0)
    s = np.select([q == 0, q == 1, q == 2], [s0, c0, -s0], -c0)
    return (c.astype(np.float32) + 1j * s.astype(np.float32)).astype(np.complex64)


def _i2f32(n: np.ndarray) -> np.ndarray:
    """Julia's Int64 -> Float32 promotion (round to nearest even)."""
    return np.asarray(n, dtype=np.int64).astype(np.float32)


def vlow32(k, l, dr: np.float32, Nf: int) -> np.ndarray:
    """``cispi(-k * l^2 * dr / Nf)`` with Julia promotion (src/zdtfdr.jl:140-142)."""
    n = -(np.asarray(k, dtype=np.int64) * np.asarray(l, dtype=np.int64) ** 2)
    return cispi32((_i2f32(n) * np.float32(dr)) / np.float32(Nf))


def prephase32(k, l, r0: np.float32, dr: np.float32, Nf: int) -> np.ndarray:
    """``cispi(k * l * (l * dr + 2 * r0) / Nf)`` (src/zdtfdr.jl:196-198)."""
    k = np.asarray(k, dtype=np.int64)
    l = np.asarray(l, dtype=np.int64)
    inner = (l.astype(np.float32) * np.float32(dr)) + np.float32(2) * np.float32(r0)
    return cispi32((_i2f32(k * l) * inner) / np.float32(Nf))


def postphase32(k, l, dr: np.float32, Nf: int) -> np.ndarray:
    """``cispi(k * l * l * dr / Nf)`` (src/zdtfdr.jl:268-270), rect window."""
    k = np.asarray(k, dtype=np.int64)
    l = np.asarray(l, dtype=np.int64)
    return cispi32((_i2f32(k * l * l) * np.float32(dr)) / np.float32(Nf))


class RefWorkspace32:
    """FP32 restatement of ``ZDTWorkspace`` and its step functions.

    Mirrors src/zdtfdr.jl:37-75 (construction: calcNl, F, Y, V, input!,
    computeV!) and the steps input!/preprocess!/convolve!/postprocess!/output!
    (src/zdtfdr.jl:209-331), in complex64 with normalized inverse FFTs.
    """

    def __init__(self, spec: np.ndarray, r0, dr, Nr: int, factors=(2, 3, 5)):
        spec = np.asarray(spec, dtype=np.float32)
        self.Nf, self.Nt = spec.shape
        self.Nk = self.Nf // 2 + 1
        self.r0 = f32(r0)
        self.dr = f32(dr)
        self.Nr = int(Nr)
        self.factors = tuple(factors)
        self.Nl = _calcNl(self.Nt, self.Nr, factors)
        self.Y = np.zeros((self.Nk, self.Nl), dtype=np.complex64)
        self.input(spec)
        self.V = self._computeV()

    # src/zdtfdr.jl:173-187
    def _computeV(self) -> np.ndarray:
        Nk, Nl, Nr, Nt, Nf = self.Nk, self.Nl, self.Nr, self.Nt, self.Nf
        V = np.zeros((Nk, Nl), dtype=np.complex64)
        k = np.arange(Nk, dtype=np.int64)[:, None]
        V[:, :Nr] = vlow32(k, np.arange(Nr)[None, :], self.dr, Nf)
        if Nt > 1:
            cols = np.arange(Nl - Nt + 1, Nl)[None, :]
            V[:, Nl - Nt + 1:] = vlow32(k, Nl - cols, self.dr, Nf)
        return sfft.fft(V, axis=1)

    # src/zdtfdr.jl:209-212
    def input(self, spec: np.ndarray) -> None:
        self.F = sfft.rfft(np.asarray(spec, dtype=np.float32), axis=0)

    # src/zdtfdr.jl:221-235
    def preprocess(self, r0=None) -> None:
        r0 = self.r0 if r0 is None else f32(r0)
        k = np.arange(self.Nk, dtype=np.int64)[:, None]
        t = np.arange(self.Nt, dtype=np.int64)[None, :]
        self.Y[:, :self.Nt] = self.F * prephase32(k, t, r0, self.dr, self.Nf)
        self.Y[:, self.Nt:] = 0

    # src/zdtfdr.jl:249-253
    def convolve(self) -> None:
        Y = sfft.fft(self.Y, axis=1)
        Y *= self.V
        self.Y = sfft.ifft(Y, axis=1)

    # src/zdtfdr.jl:297-321
    def postprocess(self, window: Window = None) -> None:
        k = np.arange(self.Nk, dtype=np.int64)[:, None]
        j = np.arange(self.Nr, dtype=np.int64)[None, :]
        pp = postphase32(k, j, self.dr, self.Nf)
        w = window64(window, self.Nf)
        Ys = self.Y[:, :self.Nr]
        if w is None:
            self.Y[:, :self.Nr] = Ys * pp
        else:
            # ComplexF32 * Float64 promotes to ComplexF64; the product with Ys
            # is formed in ComplexF64 and rounded on store into Ys.
            pp64 = pp.astype(np.complex128) * w[:, None]
            self.Y[:, :self.Nr] = (Ys.astype(np.complex128) * pp64).astype(np.complex64)

    @property
    def Ys(self) -> np.ndarray:
        return self.Y[:, :self.Nr]

    # src/zdtfdr.jl:328-331
    def output(self) -> np.ndarray:
        return sfft.irfft(self.Ys, n=self.Nf, axis=0).astype(np.float32)

    # src/zdtfdr.jl:349-362 (one batch per r0)
    def zdtfdr(self, window: Window = None, r0=None) -> np.ndarray:
        self.preprocess(r0)
        self.convolve()
        self.postprocess(window)
        return self.output()


def zdt_ref32(spec: np.ndarray, r0, dr, Nr: int, factors=(2, 3, 5),
              window: Window = None) -> np.ndarray:
    """One ZDT by the FP32 restatement (construct workspace, one batch)."""
    return RefWorkspace32(spec, r0, dr, Nr, factors).zdtfdr(window)


def unnormalized_scale(Nf: int, Nl: int) -> int:
    """Factor between unnormalized-inverse output and the normalized FDR
    (test/runtests.jl:31 divides by Nf*Nl)."""
    return Nf * Nl


# ---------------------------------------------------------------------------
# brute-force integer shifting
# ---------------------------------------------------------------------------


def _round_half_even(x) -> int:
    if isinstance(x, Fraction):
        return round(x)
    return int(round(float(x)))


def intshift(din: np.ndarray, rate) -> np.ndarray:
    """dout[:, t] = din circularly shifted so dout[m, t] = din[(m + n_t) mod Nf, t],
    n_t = round(rate*t) (src/intfdr.jl:7-14: circshift!(cout, cin, -n))."""
    din = np.asarray(din)
    Nf, Nt = din.shape
    out = np.empty_like(din)
    r = rate if isinstance(rate, Fraction) else rate
    for t in range(Nt):
        n = _round_half_even(r * t)
        out[:, t] = np.roll(din[:, t], -n)
    return out


def intfdr(spec: np.ndarray, rates: Iterable) -> np.ndarray:
    """Brute-force FDR by integer shifts and a sum over time
    (src/intfdr.jl:34-45).  Sums in float64 and rounds once."""
    spec = np.asarray(spec)
    rates = list(rates)
    Nf, Nt = spec.shape
    out = np.empty((Nf, len(rates)), dtype=spec.dtype)
    for j, r in enumerate(rates):
        out[:, j] = intshift(spec.astype(np.float64), r).sum(axis=1)
    return out


def dft_brute_fdr(spec: np.ndarray, rates: Sequence) -> np.ndarray:
    """O(Nf^2 Nt) brute-force FDR from the DFT definition (tiny inputs only):
    FDR[n, j] = (1/Nf) sum_k Herm(sum_t sum_m s[m,t] e^{-2pi i k m/Nf}
    e^{2 pi i k t r_j/Nf}) e^{2 pi i k n/Nf}, with the c2r convention that
    the imaginary part of the DC and Nyquist bins is dropped."""
    spec = np.asarray(spec, dtype=np.float64)
    Nf, Nt = spec.shape
    Nk = Nf // 2 + 1
    m = np.arange(Nf)
    out = np.empty((Nf, len(rates)))
    for j, r in enumerate(rates):
        r = float(exact(r))
        X = np.zeros(Nk, dtype=np.complex128)
        for k in range(Nk):
            acc = 0j
            for t in range(Nt):
                Fkt = np.sum(spec[:, t] * np.exp(-2j * np.pi * k * m / Nf))
                acc += Fkt * np.exp(2j * np.pi * k * t * r / Nf)
            X[k] = acc
        # Hermitian completion with c2r semantics
        full = np.zeros(Nf, dtype=np.complex128)
        full[0] = X[0].real
        for k in range(1, Nk):
            full[k] = X[k]
            full[(Nf - k) % Nf] = np.conj(X[k])
        if Nf % 2 == 0:
            full[Nf // 2] = X[Nf // 2].real
        n = np.arange(Nf)
        for nn in range(Nf):
            out[nn, j] = (np.sum(full * np.exp(2j * np.pi * np.arange(Nf) * nn / Nf)) / Nf).real
    return out
