"""Pins of the oracle's transforms (CPU).

Conventions of the reference FFTPACK wrappers (cft_buff.f:1-65,
rft_buff.f:24-86): forward e^{-i} scaled by 1/n, inverse e^{+i} unscaled;
the real transform packs c(2k+1) + i c(2k+2), k = 0..n/2, with c(2) = c(n+2)
= 0 and ignores the imaginary parts of the k = 0 and k = n/2 terms on the
inverse.  Checked against a brute-force O(n^2) DFT (closed form), against the
golden vectors of the compiled reference routines, and by round trip and
Parseval.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden", "fft_kat.npz")
NS = (6, 12, 24, 48, 64, 72, 96, 192, 384)


def dft(x, sign):
    n = len(x)
    k = np.arange(n)
    return np.exp(sign * 2j * np.pi * np.outer(k, k) / n) @ x


@pytest.mark.parametrize("n", NS)
def test_cft_closed_form(n):
    rng = np.random.default_rng(n)
    m = 2
    c = rng.standard_normal(2 * n * m)
    z = c.reshape(m, n, 2) @ np.array([1, 1j])
    f = c.copy()
    O.or_cft(n, f, 2, 2 * n, m, -1)
    F = f.reshape(m, n, 2) @ np.array([1, 1j])
    for v in range(m):
        assert np.abs(F[v] - dft(z[v], -1) / n).max() < 1e-14 * np.sqrt(n)
    b = c.copy()
    O.or_cft(n, b, 2, 2 * n, m, +1)
    B = b.reshape(m, n, 2) @ np.array([1, 1j])
    for v in range(m):
        assert np.abs(B[v] - dft(z[v], +1)).max() < 1e-13 * n
    # round trip and Parseval
    O.or_cft(n, f, 2, 2 * n, m, +1)
    assert np.abs(f - c).max() < 1e-14 * n
    assert abs(np.sum(np.abs(F) ** 2) * n - np.sum(np.abs(z) ** 2)) < 1e-12 * np.sum(np.abs(z) ** 2)


@pytest.mark.parametrize("n", NS)
def test_cft_strided(n):
    """ise/isa strides as four_to_phys uses them (FOU3D.f90:566-571)."""
    rng = np.random.default_rng(7 + n)
    igal, m = 2 * n + 6, 3
    P = rng.standard_normal(igal * n)
    a = P.copy()
    O.or_cft(n, a, igal, 2, m, +1)
    for v in range(m):
        z = P[2 * v::igal] + 1j * P[2 * v + 1::igal]
        got = a[2 * v::igal] + 1j * a[2 * v + 1::igal]
        assert np.abs(got - dft(z[:n], +1)).max() < 1e-12 * n
    assert np.array_equal(a[2 * m:igal], P[2 * m:igal])  # untouched columns


@pytest.mark.parametrize("n", NS)
def test_rft_closed_form(n):
    rng = np.random.default_rng(100 + n)
    isa, m = n + 2, 2
    x = rng.standard_normal(isa * m)
    f = x.copy()
    O.or_rft(n, f, isa, m, -1)
    for v in range(m):
        row = x[v * isa:v * isa + n]
        X = dft(row, -1)[: n // 2 + 1] / n
        got = f[v * isa:(v + 1) * isa:2] + 1j * f[v * isa + 1:(v + 1) * isa:2]
        assert np.abs(got - X).max() < 1e-15 * np.sqrt(n) * 10
        assert f[v * isa + 1] == 0.0 and f[v * isa + n + 1] == 0.0
    # inverse: sum over the Hermitian extension, Im of k=0 and k=n/2 ignored
    b = x.copy()
    O.or_rft(n, b, isa, m, +1)
    for v in range(m):
        C = x[v * isa:(v + 1) * isa:2] + 1j * x[v * isa + 1:(v + 1) * isa:2]
        C[0] = C[0].real
        C[-1] = C[-1].real
        full = np.concatenate([C, np.conj(C[-2:0:-1])])
        want = dft(full, +1).real
        assert np.abs(b[v * isa:v * isa + n] - want).max() < 1e-13 * n


@pytest.mark.skipif(not os.path.exists(GOLD), reason="golden fixtures not generated")
@pytest.mark.parametrize("n", NS)
def test_against_reference_fftpack(n):
    g = np.load(GOLD)
    for iopt in (-1, 1):
        c = g["cft_%d_%d_in" % (n, iopt)].copy()
        O.or_cft(n, c, 2, 2 * n, 3, iopt)
        ref = g["cft_%d_%d_out" % (n, iopt)]
        assert np.abs(c - ref).max() <= 1e-13 * max(1.0, n) * np.abs(ref).max()
        x = g["rft_%d_%d_in" % (n, iopt)].copy()
        O.or_rft(n, x, n + 2, 3, iopt)
        ref = g["rft_%d_%d_out" % (n, iopt)]
        if iopt > 0:  # only the n physical values are defined
            x = x.reshape(3, n + 2)[:, :n]
            ref = ref.reshape(3, n + 2)[:, :n]
        assert np.abs(x - ref).max() <= 1e-13 * max(1.0, n) * np.abs(ref).max()


def test_against_reference_live():
    # the compiled reference routines are loaded here, when the test runs, not
    # at collection (they are built only where the reference sources are)
    if O.load_ref() is None:
        pytest.skip("compiled reference routines not built")
    rng = np.random.default_rng(5)
    for n in (18, 30, 80, 120):
        c = rng.standard_normal(2 * n * 2)
        a, b = c.copy(), c.copy()
        O.ref_cft(n, a, 2, 2 * n, 2, -1)
        O.or_cft(n, b, 2, 2 * n, 2, -1)
        assert np.abs(a - b).max() < 1e-14 * n
