#!/usr/bin/env python3
"""Generate the Rys quadrature interpolation table used by the HIP kernels and the oracle.

Rys quadrature (Rys, Dupuis & King, J. Comput. Chem. 4, 154 (1983)) replaces the Boys
integral  int_0^1 P(t^2) exp(-T t^2) dt  by  sum_i w_i P(x_i)  (x_i = t_i^2) exactly for
polynomials P of degree < 2n.  The nodes/weights are the n-point Gauss rule in x = t^2 for the
measure exp(-T x) x^{-1/2}/2 dx on [0,1], whose moments are the Boys functions
F_k(T) = int_0^1 t^{2k} exp(-T t^2) dt.

Reference behaviour followed: src/lightspeed/rys.hpp:56-153 (nrys <= 9, Chebyshev bins on
[0, Tc), Gauss-Hermite/sqrt(T) asymptote for T >= Tc).  The table *values* are generated here
from first principles (none are taken from the reference):

  * exact nodes/weights: Boys moments at 60 digits (mpmath hyp1f1), the Chebyshev algorithm
    (Gautschi, "Orthogonal Polynomials: Computation and Approximation", 2004, sec. 2.1.7) for the
    Jacobi matrix, then Golub-Welsch (mpmath eigsy);
  * x = t^2 and w are fitted per uniform T-bin by degree-DEG Chebyshev interpolation and stored
    as monomial coefficients in the bin-local variable u in [-1, 1] (Horner evaluation);
  * the asymptotic switch Tc[n] is the first bin edge beyond which the Gauss-Hermite form
    x = r_i^2 / T, w = W_i / sqrt(T) is accurate to 1e-15 (absolute, relative to F_0(T)).

Output: lightspeed_amd/csrc/common/rys_table.inc  (C header, used by HIP and C code).
Run:  python lightspeed_amd/tables/gen_rys.py            (about 2-4 minutes on 8 cores)
"""
import os
import sys
import multiprocessing as mproc

import mpmath as mp
import numpy as np

NMAX = 9          # largest number of roots (L_total <= 16)
DEG = 11          # polynomial degree per bin (DEG+1 coefficients)
BIN_WIDTH = {1: 1.0, 2: 1.0, 3: 1.0, 4: 1.0, 5: 1.0, 6: 1.0, 7: 1.0, 8: 1.0, 9: 1.0}
DPS = 60

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "csrc", "common", "rys_table.inc")


def boys_moments(n2, T):
    """F_0..F_{n2-1}(T) at mp precision (downward recursion from the 1F1 closed form)."""
    T = mp.mpf(T)
    kmax = n2 - 1
    F = [mp.mpf(0)] * n2
    F[kmax] = mp.hyp1f1(kmax + mp.mpf(1) / 2, kmax + mp.mpf(3) / 2, -T) / (2 * kmax + 1)
    eT = mp.e ** (-T)
    for k in range(kmax - 1, -1, -1):
        F[k] = (2 * T * F[k + 1] + eT) / (2 * k + 1)
    return F


def rys_exact(n, T):
    """n-point Rys rule at T: (x_i, w_i) ascending in x, as mp numbers."""
    with mp.workdps(DPS):
        mom = boys_moments(2 * n, T)
        # Chebyshev algorithm: moments -> recurrence coefficients alpha_k, beta_k
        sig_prev = [mp.mpf(0)] * (2 * n)
        sig = [mom[l] for l in range(2 * n)]
        alpha = [mom[1] / mom[0]]
        beta = [mom[0]]
        for k in range(1, n):
            sig_new = [mp.mpf(0)] * (2 * n)
            for l in range(k, 2 * n - k):
                sig_new[l] = sig[l + 1] - alpha[k - 1] * sig[l] - beta[k - 1] * sig_prev[l]
            alpha.append(sig_new[k + 1] / sig_new[k] - sig[k] / sig[k - 1])
            beta.append(sig_new[k] / sig[k - 1])
            sig_prev, sig = sig, sig_new
        J = mp.matrix(n, n)
        for i in range(n):
            J[i, i] = alpha[i]
            if i + 1 < n:
                J[i, i + 1] = mp.sqrt(beta[i + 1])
                J[i + 1, i] = mp.sqrt(beta[i + 1])
        E, Q = mp.eigsy(J)
        pts = sorted([(E[i], beta[0] * Q[0, i] ** 2) for i in range(n)], key=lambda p: p[0])
        return [p[0] for p in pts], [p[1] for p in pts]


def hermite_half(n):
    """Positive nodes r_i (ascending) of the 2n-point Gauss-Hermite rule and weights for
    int_0^inf f(u^2) exp(-u^2) du = sum_i W_i f(r_i^2)."""
    with mp.workdps(DPS):
        # Golub-Welsch for Hermite weight exp(-u^2): alpha=0, beta_k = k/2
        m = 2 * n
        J = mp.matrix(m, m)
        for i in range(m - 1):
            J[i, i + 1] = mp.sqrt(mp.mpf(i + 1) / 2)
            J[i + 1, i] = J[i, i + 1]
        E, Q = mp.eigsy(J)
        pts = sorted([(E[i], mp.sqrt(mp.pi) * Q[0, i] ** 2) for i in range(m) if E[i] > 0],
                     key=lambda p: p[0])
        return [p[0] for p in pts], [p[1] for p in pts]


def cheb_nodes(deg):
    k = np.arange(deg + 1)
    return np.cos(np.pi * (k + 0.5) / (deg + 1))  # in (-1, 1)


def _bin_job(args):
    n, b, h = args
    u = cheb_nodes(DEG)
    T0 = b * h
    xs = np.zeros((n, DEG + 1))
    ws = np.zeros((n, DEG + 1))
    for j, uj in enumerate(u):
        T = mp.mpf(T0) + (mp.mpf(uj) + 1) * mp.mpf(h) / 2
        x, w = rys_exact(n, T)
        xs[:, j] = [float(v) for v in x]
        ws[:, j] = [float(v) for v in w]
    # Chebyshev interpolation -> monomial coefficients in u
    V = np.polynomial.chebyshev.chebvander(u, DEG)
    out = np.zeros((n, 2, DEG + 1))
    for i in range(n):
        for k, vals in enumerate((xs[i], ws[i])):
            cc = np.linalg.solve(V, vals)
            out[i, k, :] = np.polynomial.chebyshev.cheb2poly(cc)
    return n, b, out


def _err_job(args):
    n, T = args
    x, w = rys_exact(n, T)
    return n, T, [float(v) for v in x], [float(v) for v in w], float(boys_moments(1, T)[0])


def find_tc(n, pool):
    r, W = hermite_half(n)
    r2 = [float(v * v) for v in r]
    Wf = [float(v) for v in W]
    Ts = np.arange(10.0, 120.0, 1.0)
    res = pool.map(_err_job, [(n, float(T)) for T in Ts])
    tc = None
    for (nn, T, x, w, f0) in res:
        ex = max(abs(x[i] - r2[i] / T) for i in range(n))
        ew = max(abs(w[i] - Wf[i] / np.sqrt(T)) for i in range(n)) / f0
        if max(ex, ew) < 1e-15:
            if tc is None:
                tc = T
        else:
            tc = None
    return tc, r2, Wf


def evaluate(table, n, T):
    """Reference evaluation of the generated table (same arithmetic as the kernels)."""
    nbin, h, coefs, tc, r2, Wf = table[n]
    if T >= tc:
        return np.array(r2) / T, np.array(Wf) / np.sqrt(T)
    s = T / h
    b = int(s)
    u = 2.0 * (s - b) - 1.0
    c = coefs[b]
    x = np.zeros(n)
    w = np.zeros(n)
    for i in range(n):
        px = c[i, 0, DEG]
        pw = c[i, 1, DEG]
        for k in range(DEG - 1, -1, -1):
            px = px * u + c[i, 0, k]
            pw = pw * u + c[i, 1, k]
        x[i] = px
        w[i] = pw
    return x, w


def main():
    pool = mproc.Pool(min(8, os.cpu_count() or 1))
    table = {}
    for n in range(1, NMAX + 1):
        tc, r2, Wf = find_tc(n, pool)
        h = BIN_WIDTH[n]
        nbin = int(np.ceil(tc / h))
        tc = nbin * h
        jobs = [(n, b, h) for b in range(nbin)]
        coefs = [None] * nbin
        for (nn, b, out) in pool.map(_bin_job, jobs):
            coefs[b] = out
        table[n] = (nbin, h, coefs, tc, r2, Wf)
        print("nrys=%d Tc=%.1f nbin=%d" % (n, tc, nbin), flush=True)

    # validation at off-node points
    rng = np.random.default_rng(7)
    worst = {}
    for n in range(1, NMAX + 1):
        Ts = np.concatenate([[0.0, 1e-12, 1e-6], rng.uniform(0, table[n][3] + 5, 60)])
        res = pool.map(_err_job, [(n, float(T)) for T in Ts])
        ex_max = 0.0
        ew_max = 0.0
        for (nn, T, x, w, f0) in res:
            xa, wa = evaluate(table, n, T)
            ex_max = max(ex_max, float(np.max(np.abs(xa - np.array(x)))))
            ew_max = max(ew_max, float(np.max(np.abs(wa - np.array(w)))) / f0)
        worst[n] = (ex_max, ew_max)
        print("nrys=%d  max|dx|=%.3e  max|dw|/F0=%.3e" % (n, ex_max, ew_max), flush=True)
    pool.close()

    # emit the C header
    lines = []
    lines.append("/* Generated by lightspeed_amd/tables/gen_rys.py -- do not edit.")
    lines.append(" * Rys quadrature table: per nrys, uniform T-bins of width h on [0, Tc); per bin and root,")
    lines.append(" * DEG+1 monomial coefficients (ascending powers of u = 2*(T/h - bin) - 1) for x = t^2, then")
    lines.append(" * DEG+1 for the weight w.  Layout: [bin][root][x|w][DEG+1].  For T >= Tc:")
    lines.append(" * x_i = r2_i / T, w_i = W_i / sqrt(T) (positive half of the 2n-point Gauss-Hermite rule).")
    lines.append(" * Measured max|dx|, max|dw|/F0(T) against 60-digit values:")
    for n in range(1, NMAX + 1):
        lines.append(" *   nrys=%d: %.2e %.2e" % (n, worst[n][0], worst[n][1]))
    lines.append(" */")
    lines.append("#define LSJK_RYS_NMAX %d" % NMAX)
    lines.append("#define LSJK_RYS_DEG %d" % DEG)
    lines.append("#define LSJK_RYS_NCOEF %d" % (DEG + 1))
    off = 0
    offs = [0]
    for n in range(1, NMAX + 1):
        offs.append(off)
        off += table[n][0] * n * 2 * (DEG + 1)
    lines.append("static const int lsjk_rys_nbin[%d] = {0, %s};" % (NMAX + 1, ", ".join(str(table[n][0]) for n in range(1, NMAX + 1))))
    lines.append("static const int lsjk_rys_offset[%d] = {%s};" % (NMAX + 1, ", ".join(str(o) for o in offs)))
    lines.append("static const double lsjk_rys_h[%d] = {0.0, %s};" % (NMAX + 1, ", ".join(repr(float(table[n][1])) for n in range(1, NMAX + 1))))
    lines.append("static const double lsjk_rys_Tc[%d] = {0.0, %s};" % (NMAX + 1, ", ".join(repr(float(table[n][3])) for n in range(1, NMAX + 1))))
    # asymptotic data, packed [n][i] with offset n*(n-1)/2
    r2s = []
    Ws = []
    for n in range(1, NMAX + 1):
        r2s += table[n][4]
        Ws += table[n][5]
    lines.append("static const double lsjk_rys_asym_r2[%d] = {%s};" % (len(r2s), ", ".join("%.17e" % v for v in r2s)))
    lines.append("static const double lsjk_rys_asym_w[%d] = {%s};" % (len(Ws), ", ".join("%.17e" % v for v in Ws)))
    lines.append("static const double lsjk_rys_coef[%d] = {" % off)
    for n in range(1, NMAX + 1):
        nbin, h, coefs, tc, r2, Wf = table[n]
        for b in range(nbin):
            vals = coefs[b].reshape(-1)
            lines.append("  " + ", ".join("%.17e" % v for v in vals) + ",")
    lines.append("};")
    with open(OUT, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    print("wrote", os.path.normpath(OUT), "doubles:", off)


if __name__ == "__main__":
    sys.exit(main())
