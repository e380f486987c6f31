"""Derive the polynomial coefficients used by ``sincos_2pi_small`` in
zdt_internal.h: sin(2*pi*u) = u*P(u^2) and cos(2*pi*u) = Q(u^2) on
|u| <= 1/8, P of degree 3 and Q of degree 4, by a discrete minimax (Remez exchange on a
dense grid) for the relative error of sin and the absolute error of cos.

Run: python gen_sincos_coeffs.py   (prints the float32 coefficients and the
max error of the float32 Horner evaluation against float64).
"""
import numpy as np


def remez_even(f, weight, deg, a, iters=60):
    """Minimax fit of f(u) ~ sum_i c_i u^(2i), i<=deg, on u in [0, a]."""
    grid = np.linspace(0.0, a, 20001)
    n = deg + 2
    # initial reference: Chebyshev extrema mapped to [0, a]
    ref = a * (1 - np.cos(np.linspace(0, np.pi, n))) / 2
    for _ in range(iters):
        A = np.zeros((n, n))
        for i in range(deg + 1):
            A[:, i] = ref ** (2 * i)
        A[:, -1] = [(-1) ** m / weight(x) for m, x in enumerate(ref)]
        sol = np.linalg.solve(A, f(ref))
        c = sol[:-1]
        err = (np.polyval(c[::-1], grid ** 2) - f(grid)) * weight(grid)
        # new reference: alternating extrema
        idx = [0]
        sign = np.sign(err[0])
        for i in range(1, len(grid)):
            if np.sign(err[i]) != sign and err[i] != 0:
                idx.append(i)
                sign = np.sign(err[i])
            elif abs(err[i]) > abs(err[idx[-1]]):
                idx[-1] = i
        if len(idx) < n:
            break
        # keep the n largest consecutive alternating points
        idx = np.array(idx)
        while len(idx) > n:
            if abs(err[idx[0]]) < abs(err[idx[-1]]):
                idx = idx[1:]
            else:
                idx = idx[:-1]
        ref = grid[idx]
    return c, np.max(np.abs(err))


def main():
    a = 0.125
    ps, es = remez_even(lambda u: np.where(u == 0, 2 * np.pi, np.sin(2 * np.pi * u) / np.where(u == 0, 1, u)),
                        lambda u: np.ones_like(u), 3, a)
    pc, ec = remez_even(lambda u: np.cos(2 * np.pi * u), lambda u: np.ones_like(u), 4, a)
    ps32 = ps.astype(np.float32)
    pc32 = pc.astype(np.float32)
    u = np.linspace(-a, a, 200001).astype(np.float32)
    z = (u * u).astype(np.float32)

    def horner(c):
        acc = np.float32(c[-1])
        for ci in c[-2::-1]:
            acc = (acc.astype(np.float64) * z + np.float64(ci)).astype(np.float32)
        return acc

    s = (u.astype(np.float64) * horner(ps32)).astype(np.float32)
    c = horner(pc32)
    print("sin coeffs (z^3..z^0):", ", ".join(f"{x:.8e}" for x in ps32[::-1]))
    print("cos coeffs (z^4..z^0):", ", ".join(f"{x:.8e}" for x in pc32[::-1]))
    print("remez err sin/u, cos:", es, ec)
    print("float32 max abs err sin:", np.max(np.abs(s - np.sin(2 * np.pi * u.astype(np.float64)))))
    print("float32 max abs err cos:", np.max(np.abs(c - np.cos(2 * np.pi * u.astype(np.float64)))))


if __name__ == "__main__":
    main()
