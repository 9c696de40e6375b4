"""Generate rys_table.h: piecewise-polynomial Rys roots and weights.

The Rys rule of order n for argument X is the n-point Gauss rule of the
measure exp(-X u) u^(-1/2)/2 du on u in [0,1], i.e.
    sum_i w_i f(u_i) = int_0^1 f(t^2) exp(-X t^2) dt   for deg f < 2n,
with moments F_k(X) (the Boys function).  Here each rule is computed in
40-digit arithmetic (mpmath): moments by their power series, recurrence
coefficients by Chebyshev's algorithm, nodes and weights from the
eigen-decomposition of the Jacobi matrix (Golub-Welsch).

Tabulation: X in [0, XA_n) is cut into intervals of width H; on each, every
root u_i and weight w_i is a degree-DEG polynomial in s = 2(X-X_k)/H - 1,
fitted at Chebyshev points and stored as monomial coefficients (Horner on
the device).  For X >= XA_n the large-X limit is used:
    u_i = s_i / X,  w_i = l_i / sqrt(X),
with (s_i, 2 l_i) the Gauss rule of s^(-1/2) e^(-s) on [0, inf); XA_n is
chosen so that this limit is accurate to < 1e-15 relative.

FP32 table (rys_table32.h, `--fp32`): the same construction with degree
DEG32 = 6 on unit intervals (max relative error ~2e-8 over n <= 5, below
float rounding), stored as float with each function padded to 8 entries so
that a function's coefficients are two aligned 16-byte loads; the large-X
limit (evaluated in FP64 on the device) takes over where it is accurate to
< 5e-8 relative.

Run: python gen_rys_table.py [--fp32]  (writes rys_table.h or rys_table32.h
next to this file).
"""
import multiprocessing as mpr
import os
import sys

import mpmath as mp

mp.mp.dps = 40
NMAX = 5
H = 1.0
DEG = 11
HERE = os.path.dirname(os.path.abspath(__file__))


def moments(X, K):
    X = mp.mpf(X)
    out = []
    for k in range(K):
        term = 1 / mp.mpf(2 * k + 1)
        s = term
        i = 1
        while True:
            term *= 2 * X / (2 * k + 2 * i + 1)
            s += term
            i += 1
            if term < s * mp.mpf(10) ** (-45):
                break
        out.append(mp.e ** (-X) * s)
    return out


def gauss_from_moments(mu, n):
    """Chebyshev's algorithm + Golub-Welsch; returns sorted nodes, weights."""
    a = [0] * n
    b = [0] * n
    sig_prev = [mp.mpf(0)] * (2 * n)
    sig = list(mu)
    a[0] = mu[1] / mu[0]
    b[0] = mu[0]
    for k in range(1, n):
        sig_new = [mp.mpf(0)] * (2 * n)
        for l in range(k, 2 * n - k):
            sig_new[l] = sig[l + 1] - a[k - 1] * sig[l] - b[k - 1] * sig_prev[l]
        a[k] = sig_new[k + 1] / sig_new[k] - sig[k] / sig[k - 1]
        b[k] = sig_new[k] / sig[k - 1]
        sig_prev, sig = sig, sig_new
    J = mp.matrix(n, n)
    for i in range(n):
        J[i, i] = a[i]
        if i + 1 < n:
            J[i, i + 1] = J[i + 1, i] = mp.sqrt(b[i + 1])
    E, Q = mp.eigsy(J)
    order = sorted(range(n), key=lambda i: E[i])
    return [E[i] for i in order], [b[0] * Q[0, i] ** 2 for i in order]


def rys(n, X):
    return gauss_from_moments(moments(X, 2 * n), n)


def laguerre_half(n):
    """Gauss rule for s^(-1/2) e^(-s) on [0, inf): moments Gamma(k + 1/2)."""
    mu = [mp.gamma(k + mp.mpf(1) / 2) for k in range(2 * n)]
    return gauss_from_moments(mu, n)


def asym(n, X):
    s, lam = laguerre_half(n)
    X = mp.mpf(X)
    return [si / X for si in s], [li / (2 * mp.sqrt(X)) for li in lam]


DEG32 = 6
STRIDE32 = 8


def fit_interval(args):
    n, k = args[:2]
    deg = args[2] if len(args) > 2 else DEG
    return _fit(n, k, deg)


def _fit(n, k, DEG):
    x0 = k * H
    nodes = [mp.cos(mp.pi * (j + mp.mpf(1) / 2) / (DEG + 1)) for j in range(DEG + 1)]
    vals = []
    for s in nodes:
        u, w = rys(n, x0 + (s + 1) * H / 2)
        vals.append(u + w)
    coefs = []
    for f in range(2 * n):
        # Chebyshev coefficients c_m, then monomial coefficients in s
        c = []
        for m in range(DEG + 1):
            acc = mp.mpf(0)
            for j, s in enumerate(nodes):
                acc += vals[j][f] * mp.cos(m * mp.acos(s))
            c.append(acc * (1 if m == 0 else 2) / (DEG + 1))
        T = [[mp.mpf(1)] + [mp.mpf(0)] * DEG, [mp.mpf(0), mp.mpf(1)] + [mp.mpf(0)] * (DEG - 1)]
        for m in range(2, DEG + 1):
            T.append([2 * (T[m - 1][i - 1] if i > 0 else 0) - T[m - 2][i] for i in range(DEG + 1)])
        mono = [sum(c[m] * T[m][i] for m in range(DEG + 1)) for i in range(DEG + 1)]
        coefs.append([float(v) for v in mono])
    return n, k, coefs


def threshold(n, tol=1e-15, start=10.0):
    """Smallest multiple of H (from start) past which the large-X limit is
    < tol relative."""
    X = start
    while True:
        u, w = rys(n, X)
        ua, wa = asym(n, X)
        err = max(max(abs((a - b) / b) for a, b in zip(ua, u)),
                  max(abs((a - b) / b) for a, b in zip(wa, w)))
        if err < tol:
            return X
        X += H


def main32():
    xa = {n: threshold(n, 5e-8, 4.0) for n in range(1, NMAX + 1)}
    jobs = [(n, k, DEG32) for n in range(1, NMAX + 1) for k in range(int(round(xa[n] / H)))]
    with mpr.Pool(min(8, os.cpu_count() or 1)) as pool:
        res = pool.map(fit_interval, jobs, chunksize=2)
    table = {(n, k): coefs for n, k, coefs in res}
    nint = {n: int(round(xa[n] / H)) for n in range(1, NMAX + 1)}
    offs = [0]
    for n in range(1, NMAX + 1):
        offs.append(offs[-1] + nint[n] * 2 * n * STRIDE32)
    lines = ["// Generated by gen_rys_table.py --fp32 -- do not edit.",
             "// FP32 piecewise Rys roots u_i = t_i^2 and weights w_i, n = 1..%d." % NMAX,
             "#pragma once", "",
             "#ifndef RYS_DEV",
             "#define RYS_DEV __device__",
             "#endif", "",
             "#define RYS_DEG32 %d" % DEG32,
             "#define RYS_STRIDE32 %d" % STRIDE32, ""]
    lines.append("RYS_DEV const int rys_off32[%d] = {%s};" % (NMAX, ", ".join(str(o) for o in offs[:NMAX])))
    lines.append("RYS_DEV const double rys_xasym32[%d] = {0, %s};" %
                 (NMAX + 1, ", ".join("%.1f" % xa[n] for n in range(1, NMAX + 1))))
    lines.append("// coefficients: [n][interval][function: u_1..u_n, w_1..w_n][power 0..DEG32, zero pad]")
    lines.append("RYS_DEV const float __attribute__((aligned(16))) rys_coef32[%d] = {" % offs[-1])
    body = []
    for n in range(1, NMAX + 1):
        for k in range(nint[n]):
            for f in table[(n, k)]:
                v = list(f) + [0.0] * (STRIDE32 - len(f))
                body.append("  " + ", ".join("%.9ef" % x for x in v))
    lines.append(",\n".join(body) + "};")
    with open(os.path.join(HERE, "rys_table32.h"), "w") as f:
        f.write("\n".join(lines) + "\n")
    print("fp32 thresholds", xa, "floats", offs[-1])


def main():
    xa = {n: threshold(n) for n in range(1, NMAX + 1)}
    jobs = [(n, k) for n in range(1, NMAX + 1) for k in range(int(round(xa[n] / H)))]
    with mpr.Pool(min(8, os.cpu_count() or 1)) as pool:
        res = pool.map(fit_interval, jobs, chunksize=4)
    table = {}
    for n, k, coefs in res:
        table[(n, k)] = coefs
    lines = ["// Generated by gen_rys_table.py -- do not edit.",
             "// Piecewise Rys roots u_i = t_i^2 and weights w_i, n = 1..%d." % NMAX,
             "#pragma once", "",
             "#ifndef RYS_DEV",
             "#define RYS_DEV __device__",
             "#endif", "",
             "#define RYS_NMAX %d" % NMAX,
             "#define RYS_H %.17g" % H,
             "#define RYS_INV_H %.17g" % (1.0 / H),
             "#define RYS_DEG %d" % DEG, ""]
    offs = [0]
    for n in range(1, NMAX + 1):
        offs.append(offs[-1] + int(round(xa[n] / H)) * 2 * n * (DEG + 1))
    lines.append("RYS_DEV const int rys_nint[RYS_NMAX + 1] = {0, %s};" %
                 ", ".join(str(int(round(xa[n] / H))) for n in range(1, NMAX + 1)))
    lines.append("RYS_DEV const int rys_off[RYS_NMAX + 1] = {%s};" %
                 ", ".join(str(o) for o in offs[:NMAX + 1]))
    lines.append("RYS_DEV const double rys_xasym[RYS_NMAX + 1] = {0, %s};" %
                 ", ".join("%.1f" % xa[n] for n in range(1, NMAX + 1)))
    sl = []
    for n in range(1, NMAX + 1):
        s, lam = laguerre_half(n)
        sl.append("  {%s}" % ", ".join("%.17e" % float(v) for v in s + [li / 2 for li in lam]))
    lines.append("// large-X limit: u_i = s_i/X, w_i = l_i/sqrt(X); rows hold s_1..s_n, l_1..l_n")
    lines.append("RYS_DEV const double rys_asym[RYS_NMAX][2 * RYS_NMAX] = {")
    lines.append(",\n".join(sl) + "};")
    lines.append("// coefficients: [n][interval][function: u_1..u_n, w_1..w_n][power 0..DEG]")
    lines.append("RYS_DEV const double rys_coef[%d] = {" % offs[-1])
    body = []
    for n in range(1, NMAX + 1):
        for k in range(int(round(xa[n] / H))):
            for f in table[(n, k)]:
                body.append("  " + ", ".join("%.17e" % v for v in f))
    lines.append(",\n".join(body) + "};")
    with open(os.path.join(HERE, "rys_table.h"), "w") as f:
        f.write("\n".join(lines) + "\n")
    print("thresholds", xa, "doubles", offs[-1])


if __name__ == "__main__":
    sys.exit(main32() if "--fp32" in sys.argv[1:] else main())
