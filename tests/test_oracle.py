"""Pins of the CPU oracle (no GPU).

Each oracle function is checked against something that does not restate its
own formula: the reference's known answers (test/runtests.jl:5-31), closed
forms (integer-shift sums), brute-force DFT sums on tiny inputs, an
independent algorithm (Bluestein vs definition), symmetries, and hand-derived
values for the host arithmetic.
"""
import json
from fractions import Fraction
from pathlib import Path

import numpy as np
import pytest

from oracle import host_oracle as H
from oracle import synth
from oracle import zdt_oracle as Z

GOLD = json.loads((Path(__file__).parent / "golden" / "runtests_3x2.json").read_text())


def _d():
    return np.asarray(GOLD["d"], dtype=np.float32)


def _rates():
    r = GOLD["rates"]
    return [r["first"] + i * r["step"] for i in range(r["length"])]


def normwise(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


# ------------------------------------------------------------ known answers

def test_runtests_intshift_intfdr():
    d = _d()
    assert np.array_equal(Z.intshift(d, GOLD["intshift_rate"]), np.asarray(GOLD["dshift1_expected"], np.float32))
    assert np.array_equal(Z.intfdr(d, _rates()), np.asarray(GOLD["fdr_expected"], np.float32))


def test_runtests_fftfdr_fdshift():
    d = _d()
    np.testing.assert_allclose(Z.fdshift64(d, 1), GOLD["dshift1_expected"], atol=1e-12)
    np.testing.assert_allclose(Z.fftfdr64(d, _rates()), GOLD["fdr_expected"], atol=1e-12)


@pytest.mark.parametrize("fn", ["fdr64", "zdt64", "zdt_ref32"])
def test_runtests_zdt(fn):
    d = _d()
    r = GOLD["rates"]
    out = getattr(Z, fn)(d, r["first"], r["step"], r["length"])
    np.testing.assert_allclose(out, GOLD["fdr_expected"], atol=3e-7)


def test_runtests_unnormalized_convention():
    # test/runtests.jl:31 divides by Nf*Nl = 12; the normalized oracle times
    # that factor is what an unnormalized inverse produces
    assert H.calcNl(2, 3) == GOLD["Nl"]
    assert Z.unnormalized_scale(3, GOLD["Nl"]) == GOLD["zdtfdr_scale_in_test"]


# ------------------------------------------------------------ closed forms

@pytest.mark.parametrize("Nf", [64, 63, 250])
def test_fdr64_equals_intfdr_at_integer_rates(Nf):
    rng = np.random.default_rng(Nf)
    s = rng.exponential(1.0, size=(Nf, 7)).astype(np.float32)
    rates = list(range(-3, 4))
    ref = Z.intfdr(s.astype(np.float64), rates)
    out = Z.fdr64(s, -3, 1, 7)
    assert normwise(out, ref) < 1e-12


def test_intfdr_half_integer_rounding():
    # round(r*t) is round-half-even (src/intfdr.jl:10): rate 0.5 shifts t=1 by 0, t=3 by 2
    s = np.zeros((8, 4), np.float32)
    s[0, :] = 1
    out = Z.intshift(s, 0.5)
    assert out[0, 1] == 1 and out[0, 0] == 1      # n = round(0.5) = 0
    assert out[(0 - 1) % 8, 2] == 1               # n = round(1.0) = 1 -> din[m+1]
    assert out[(0 - 2) % 8, 3] == 1               # n = round(1.5) = 2


@pytest.mark.parametrize("Nf,Nt", [(16, 3), (15, 4), (10, 2)])
def test_fdr64_brute_force_dft(Nf, Nt):
    rng = np.random.default_rng(7 + Nf)
    s = rng.normal(size=(Nf, Nt)).astype(np.float32)
    rates = [Fraction(-3, 4), Fraction(1, 3), Fraction(5, 2)]
    brute = Z.dft_brute_fdr(s, rates)
    out = Z.fdr64(s, None, rates=rates)
    np.testing.assert_allclose(out, brute, atol=1e-11)


# ------------------------------------------------------------ independent algorithms

@pytest.mark.parametrize("Nf,Nt,r0,dr,Nr,factors", [
    (256, 8, -3.0, 0.25, 25, (2, 3, 5)),
    (255, 5, -1.5, 0.125, 9, (2, 3, 5)),
    (128, 16, -4.0, 0.0625, 129, (2,)),
    (128, 16, -4.0, 0.0625, 129, (2, 3, 5, 7)),
    (100, 1, 0.5, 0.5, 3, (2, 3, 5)),
])
def test_zdt64_equals_definition(Nf, Nt, r0, dr, Nr, factors):
    s = synth.spectrogram(Nf, Nt, [(17.2, 1.5, 20.0)], seed=Nf + Nt)
    a = Z.zdt64(s, r0, dr, Nr, factors)
    b = Z.fdr64(s, r0, dr, Nr)
    assert normwise(a, b) < 1e-12


def test_ref32_close_to_fp64():
    s = synth.spectrogram(4096, 16, [(100.3, 2.25, 20.0), (3000, -1.5, 10.0)], seed=5)
    a = Z.zdt_ref32(s, -4, 0.0625, 129)
    b = Z.fdr64(s, -4, 0.0625, 129)
    # the reference's Float32 phase arguments bound its distance to FP64
    assert normwise(a, b) < 5e-6


# ---- the reference's Float32 phase arguments, derived by hand
#
# Julia evaluates `-k * l^2 * δr / Nf` (src/zdtfdr.jl:141) left to right:
# the Int64 product -k*l^2, promoted to Float32 (round to nearest even) for
# the product with the Float32 δr (rounded), then divided by Float32(Nf)
# (exact for a power of two).  `k * l * (l * δr + 2 * r0) / Nf` (:197) and
# `k * l * l * δr / Nf` (:269) likewise.  rne24 below rounds an exact
# rational to Float32 with integer arithmetic only (no numpy float32).

def rne24(x: Fraction) -> Fraction:
    if x == 0:
        return Fraction(0)
    sgn = -1 if x < 0 else 1
    x = abs(x)
    e = x.numerator.bit_length() - x.denominator.bit_length()
    if Fraction(2) ** e > x:
        e -= 1
    if Fraction(2) ** (e + 1) <= x:
        e += 1
    q = x / Fraction(2) ** (e - 23)  # in [2^23, 2^24)
    m = q.numerator // q.denominator
    rem = q - m
    if rem > Fraction(1, 2) or (rem == Fraction(1, 2) and m % 2 == 1):
        m += 1
    return sgn * m * Fraction(2) ** (e - 23)


def wrap2(x: Fraction) -> float:
    """x mod 2 into [-1, 1): the cispi argument's angle over pi."""
    y = x - 2 * ((x + 1) // 2)
    return float(y)


def angle_over_pi(z) -> float:
    return float(np.angle(np.complex128(z)) / np.pi)


def test_rne24_hand_values():
    assert rne24(Fraction(2 ** 24 + 1)) == 2 ** 24          # tie to even
    assert rne24(Fraction(2 ** 24 + 3)) == 2 ** 24 + 4      # tie to even (up)
    assert rne24(Fraction(2 ** 25 + 3)) == 2 ** 25 + 4      # above the tie
    assert rne24(Fraction(1, 3)) == Fraction(11184811, 2 ** 25)


# (k, l, discriminating): where the Float32 argument differs from the exact
# one by more than twice the check's tolerance, so an exact-phase oracle fails
@pytest.mark.parametrize("k,l,disc", [(524287, 127, True), (524289, 128, False), (300001, 100, False),
                                      (65535, 143, False), (499999, 111, True)])
def test_vlow32_argument(k, l, disc):
    dr, Nf = Fraction(1, 16), 1 << 20
    arg32 = rne24(rne24(Fraction(-k * l * l)) * dr) / Nf
    exact = Fraction(-k * l * l) * dr / Nf
    got = angle_over_pi(Z.vlow32(k, l, np.float32(0.0625), Nf))
    assert abs(got - wrap2(arg32)) < 2e-6
    assert (abs(wrap2(arg32) - wrap2(exact)) > 4e-6) == disc


@pytest.mark.parametrize("k,l", [(524287, 15), (400001, 13), (123457, 9)])
def test_prephase32_argument(k, l):
    r0, dr, Nf = Fraction(-4), Fraction(1, 16), 1 << 20
    inner = rne24(rne24(Fraction(l)) * dr + 2 * r0)
    arg32 = rne24(rne24(Fraction(k * l)) * inner) / Nf
    got = angle_over_pi(Z.prephase32(k, l, np.float32(-4.0), np.float32(0.0625), Nf))
    assert abs(got - wrap2(arg32)) < 2e-6
    # k*l < 2^24 is exact in Float32; the rounding of the product with the
    # inner factor is what the reference adds (a few 1e-7 of pi here)
    exact = Fraction(k * l) * (l * dr + 2 * r0) / Nf
    assert abs(wrap2(arg32) - wrap2(exact)) < 1e-5


@pytest.mark.parametrize("k,l", [(524287, 127), (500001, 77)])
def test_postphase32_argument(k, l):
    dr, Nf = Fraction(1, 16), 1 << 20
    arg32 = rne24(rne24(Fraction(k * l * l)) * dr) / Nf
    got = angle_over_pi(Z.postphase32(k, l, np.float32(0.0625), Nf))
    assert abs(got - wrap2(arg32)) < 2e-6
    exact = Fraction(k * l * l) * dr / Nf
    assert abs(wrap2(arg32) - wrap2(exact)) > 4e-6


def test_ref32_distance_to_fp64_at_full_size():
    """At 2^20 x 16 x 129 the reference's Float32 phase arguments put its
    FP32 output 8e-6 (normwise) from the FP64 definition; FP32 FFTs with
    exact phases land near 1e-7.  The bracket pins the restatement to the
    reference's rounding: an exact-phase FP32 version fails it."""
    s = synth.cfg2(seed=1234)
    a = Z.zdt_ref32(s, -4.0, 0.0625, 129)
    # two columns of the FP64 definition suffice for the scale of the error:
    # the rounding grows with k*l^2, i.e. with the rate index
    cols = [0, 128]
    rates = [Fraction(-4) + Fraction(j, 16) for j in cols]
    b = Z.fdr64(s, None, rates=rates)
    err = float(np.max(np.abs(a[:, cols] - b)) / np.max(np.abs(b)))
    assert 4e-6 <= err <= 1.2e-5, err


# ------------------------------------------------------------ invariants

def test_time_reversal_symmetry():
    # mirroring channels n -> -n mirrors rates r -> -r
    s = synth.spectrogram(128, 6, [(30.0, 1.25, 20.0)], seed=3)
    sm = np.asfortranarray(s[(-np.arange(128)) % 128, :])
    a = Z.fdr64(s, -2, 0.25, 17)
    b = Z.fdr64(sm, 2, -0.25, 17)
    np.testing.assert_allclose(a, b[(-np.arange(128)) % 128, :], atol=1e-10)


def test_linearity():
    rng = np.random.default_rng(11)
    s1 = rng.normal(size=(96, 5)).astype(np.float32)
    s2 = rng.normal(size=(96, 5)).astype(np.float32)
    a = Z.fdr64(2 * s1.astype(np.float64) - 3 * s2.astype(np.float64), -1, 0.25, 9)
    b = 2 * Z.fdr64(s1, -1, 0.25, 9) - 3 * Z.fdr64(s2, -1, 0.25, 9)
    np.testing.assert_allclose(a, b, atol=1e-10)


@pytest.mark.parametrize("Nf", [64, 65])
def test_hamming_is_channel_convolution(Nf):
    # w(k) = 0.53836 + 0.46164 cos(2 pi k/Nf) multiplies the spectrum, i.e.
    # convolves each FDR column with [0.23082, 0.53836, 0.23082]
    s = synth.spectrogram(Nf, 4, [(10.0, 0.75, 5.0)], seed=2)
    a = Z.fdr64(s, -1, 0.5, 5, window="hamming")
    b = Z.fdr64(s, -1, 0.5, 5)
    conv = 0.53836 * b + 0.23082 * (np.roll(b, 1, axis=0) + np.roll(b, -1, axis=0))
    np.testing.assert_allclose(a, conv, atol=1e-10)


def test_custom_window_equals_named():
    s = synth.spectrogram(64, 4, [(10.0, 0.75, 5.0)], seed=2)
    a = Z.fdr64(s, -1, 0.5, 5, window=lambda n, N: 1)
    b = Z.fdr64(s, -1, 0.5, 5)
    np.testing.assert_array_equal(a, b)


def test_ref32_stepwise_hamming_matches_fp64():
    s = synth.spectrogram(512, 8, [(100.0, 1.0, 20.0)], seed=9)
    w = Z.RefWorkspace32(s, -2, 0.25, 17)
    a = w.zdtfdr("hamming")
    b = Z.fdr64(s, -2, 0.25, 17, window="hamming")
    assert normwise(a, b) < 1e-6


def test_window_errors():
    with pytest.raises(ValueError, match="unsupported window type"):
        Z.window64("blackman", 16)


# ------------------------------------------------------------ host arithmetic

def test_calcNl_values():
    assert H.calcNl(2, 3) == 4
    assert H.calcNl(16, 129) == 144
    assert H.calcNl(16, 33) == 48
    assert H.calcNl(32, 257) == 288
    assert H.calcNl(60, 481) == 540
    assert H.calcNl(64, 1025) == 1125
    assert H.calcNl(16, 129, (2,)) == 256
    assert H.growNr(16, 129) == 129
    assert H.growNr(16, 100) == 105  # nextprod(115) = 120 = 2^3*3*5 -> 120-16+1
    for n in range(1, 400):
        m = H.nextprod((2, 3, 5), n)
        assert m >= n
        assert all(not H._smooth(x, (2, 3, 5)) for x in range(n, m))


def test_estimate_memory_hand_value():
    # Nf=1024, Nt=16, Nr=129: Nl=144; 4*(1024*16*4 + 1024*144*3 + 1024*129*1)
    assert H.estimate_memory(1024, 16, 129) == 4 * (1024 * 16 * 4 + 1024 * 144 * 3 + 1024 * 129)


def test_batchrates_hand_cases():
    b = H.batchrates(16, 0.1, 4.0, Nrb=100)
    assert b == [(Fraction(-8, 3), Fraction(1, 15), 81)]
    b = H.batchrates(16, 0.1, 8.0, Nrb=50)
    assert [x[2] for x in b] == [57, 57, 57]
    assert b[0][0] == Fraction(-85, 15) and b[0][1] == Fraction(1, 15)
    assert b[-1][0] + 56 * b[-1][1] == Fraction(85, 15)
    b = H.batchrates(16, -0.1, 8.0, Nrb=50)
    assert b[0][0] == Fraction(85, 15) and b[0][1] == Fraction(-1, 15)
    # odd bonus: nc = -5..5, Nrspan=11, Nrb=growNr(4,11)=nextprod(14)-3=12 ->
    # bonus 1 -> nc1 -= round(0.5, RoundFromZero)=1, nc2 += trunc(0.5)=0
    b = H.batchrates(4, 1.0, 5.0, Nrb=100)
    assert b == [(Fraction(-6, 3), Fraction(1, 3), 12)]


def test_fdrstats_hits_hand_case():
    f = np.array([[1, 2], [3, 4], [5, 9]], dtype=np.float32)
    m, s = H.fdrstats(f)
    assert m == 3.0 and s == 2.0
    assert H.fdrdenormalize(1, m, s) == 5.0
    hits = H.findprotohits(f, 1, snr=True)
    assert hits.tolist() == [[3, 1], [3, 2]]
    hits2 = H.findprotohits([f, f], 5.0)
    assert hits2.tolist() == [[3, 1], [3, 2], [3, 3], [3, 4]]
    n = H.fdrnormalize(f)
    np.testing.assert_allclose(n, (f - 3.0) / 2.0)
