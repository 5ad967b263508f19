"""The generated Rys and Gauss-Hermite rules (CPU only).

Pinned by what the mathematics fixes, independently of the generator:
  * moments: sum_i w_i x_i^k = F_k(T) for k < 2n (Boys functions at 30 digits, mpmath quad);
  * n = 1 closed form: w = F_0(T), x = F_1(T) / F_0(T);
  * T = 0: Gauss-Legendre on [-1, 1] (positive half, weights halved) from numpy;
  * large T: the 2n-point Gauss-Hermite rule, x = r^2 / T, w = W / sqrt(T);
  * Gauss-Hermite rules against numpy.polynomial.hermite.hermgauss.
A wrong node ordering, a dropped root or a mis-scaled weight fails the moment test.
"""
import os
import re
import sys

import mpmath as mp
import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle as O  # noqa: E402


def boys(k, T):
    with mp.workdps(30):
        return float(mp.quad(lambda t: t ** (2 * k) * mp.e ** (-mp.mpf(T) * t * t), [0, 1]))


@pytest.mark.parametrize("n", range(1, 10))
def test_rys_moments(n):
    rng = np.random.default_rng(100 + n)
    Ts = np.concatenate([[0.0, 1.0e-9, 0.5, 7.3], rng.uniform(0, 90, 8)])
    x, w = O.rys(n, Ts)
    for i, T in enumerate(Ts):
        for k in range(2 * n):
            F = boys(k, T)
            assert abs(np.sum(w[i] * x[i] ** k) - F) <= 2e-14 * max(1.0, F) + 1e-15, (n, T, k)
        assert np.all(np.diff(x[i]) > 0) and np.all(x[i] > 0) and np.all(x[i] < 1)


def test_rys_one_root_closed_form():
    Ts = np.linspace(0, 60, 97)
    x, w = O.rys(1, Ts)
    for i, T in enumerate(Ts):
        F0, F1 = boys(0, T), boys(1, T)
        assert abs(w[i, 0] - F0) < 1e-15 * 20
        assert abs(x[i, 0] - F1 / F0) < 1e-14


@pytest.mark.parametrize("n", [1, 2, 3, 5, 9])
def test_rys_limits(n):
    # T = 0: positive Gauss-Legendre nodes squared, weights of the half interval
    t, W = np.polynomial.legendre.leggauss(2 * n)
    x0, w0 = O.rys(n, np.array([0.0]))
    assert np.allclose(x0[0], np.sort(t[t > 0] ** 2), atol=1e-14)
    assert np.allclose(w0[0], W[t > 0][np.argsort(t[t > 0])], atol=1e-14)
    # large T: Gauss-Hermite
    r, H = np.polynomial.hermite.hermgauss(2 * n)
    T = 200.0
    xi, wi = O.rys(n, np.array([T]))
    order = np.argsort(r[r > 0])
    assert np.allclose(xi[0], (r[r > 0] ** 2)[order] / T, rtol=1e-13)
    assert np.allclose(wi[0], H[r > 0][order] / np.sqrt(T), rtol=1e-13)


def test_gh_table_against_numpy():
    src = open(os.path.join(ROOT, "lightspeed_amd", "csrc", "common", "gh_table.inc")).read()
    xs = np.array([float(v) for v in re.search(r"lsjk_gh_x\[\d+\] = \{(.*?)\}", src, re.S).group(1).split(",")])
    ws = np.array([float(v) for v in re.search(r"lsjk_gh_w\[\d+\] = \{(.*?)\}", src, re.S).group(1).split(",")])
    for n in range(1, 21):
        o = n * (n - 1) // 2
        X, W = np.polynomial.hermite.hermgauss(n)
        assert np.allclose(xs[o:o + n], X, atol=2e-15)
        assert np.allclose(ws[o:o + n], W, rtol=1e-13, atol=1e-300)
