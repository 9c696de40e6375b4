"""Generate boys_table.h: the Boys function F_m(T) on a grid, for the
Obara-Saika / Head-Gordon-Pople J/K kernels of the light classes.

    F_m(T) = int_0^1 t^(2m) exp(-T t^2) dt

is tabulated at T_g = g / BOYS_INV_H, g = 0 .. BOYS_NG - 1, for
m = 0 .. BOYS_MTAB - 1 (row g holds F_0(T_g) .. F_{MTAB-1}(T_g), padded to a
16-byte multiple).  On the device, F_m(T) for T below BOYS_TASYM comes from
the Taylor series about the nearest grid point, using dF_m/dT = -F_{m+1}:

    F_m(T_g + d) = sum_{k=0..BOYS_K} F_{m+k}(T_g) (-d)^k / k!,  |d| <= H/2,

whose truncation error is below (H/2)^(K+1)/(K+1)! relative (F_{m+K+1} <=
F_m): 2.9e-18 for H = 1/4, K = 10.  The table (21.6 KB) is small enough to
be staged in LDS by the J/K kernels that use it.  At and above BOYS_TASYM the asymptotic
form F_m(T) = (2m-1)!! / 2^(m+1) sqrt(pi / T^(2m+1)) holds to < 2e-16
relative for m <= BOYS_MTAB - BOYS_K - 1.

Each value is computed in 40-digit arithmetic (mpmath) from the confluent
hypergeometric form F_m(T) = 1F1(m+1/2; m+3/2; -T) / (2m+1) and rounded to
double.

Run: python gen_boys_table.py  (writes boys_table.h next to this file).
"""
import os

import mpmath as mp

mp.mp.dps = 40
INV_H = 4
TASYM = 48
K = 10
MTAB = 14        # m = 0 .. 13: classes up to L = MTAB - K - 1 = 3
STRIDE = 14      # doubles per row (112 bytes; consecutive rows 28 banks apart in LDS)
HERE = os.path.dirname(os.path.abspath(__file__))


def boys(m, T):
    T = mp.mpf(T)
    if T == 0:
        return 1 / mp.mpf(2 * m + 1)
    return mp.hyp1f1(m + mp.mpf(1) / 2, m + mp.mpf(3) / 2, -T) / (2 * m + 1)


def main():
    ng = TASYM * INV_H + 1
    rows = []
    for g in range(ng):
        T = mp.mpf(g) / INV_H
        rows.append([boys(m, T) for m in range(MTAB)] + [mp.mpf(0)] * (STRIDE - MTAB))
    # check the asymptotic form at TASYM for the orders the Taylor scheme serves
    worst = mp.mpf(0)
    for m in range(MTAB - K):
        for T in (TASYM, TASYM + 0.5, TASYM + 3, 80):
            ex = boys(m, T)
            dfac = mp.fac2(2 * m - 1) if m > 0 else 1
            asym = dfac / mp.mpf(2) ** (m + 1) * mp.sqrt(mp.pi / mp.mpf(T) ** (2 * m + 1))
            worst = max(worst, abs(asym - ex) / ex)
    assert worst < 2e-16, worst
    out = []
    out.append("// Generated by gen_boys_table.py -- do not edit.")
    out.append("// Boys function F_m(T) on the grid T_g = g / BOYS_INV_H (row g: m = 0 .. BOYS_MTAB-1).")
    out.append("#pragma once")
    out.append("")
    out.append("#ifndef BOYS_DEV")
    out.append("#define BOYS_DEV __device__")
    out.append("#endif")
    out.append("")
    out.append("#define BOYS_INV_H %d" % INV_H)
    out.append("#define BOYS_TASYM %d" % TASYM)
    out.append("#define BOYS_K %d" % K)
    out.append("#define BOYS_MTAB %d" % MTAB)
    out.append("#define BOYS_STRIDE %d" % STRIDE)
    out.append("#define BOYS_NG %d" % ng)
    out.append("// asymptotic form accurate to %.1e relative at and above BOYS_TASYM" % float(worst))
    out.append("")
    out.append("BOYS_DEV const double boys_grid[%d] __attribute__((aligned(16))) = {" % (ng * STRIDE))
    for r in rows:
        out.append("  " + ", ".join(mp.nstr(v, 17, min_fixed=1, max_fixed=0) if v != 0 else "0.0" for v in r) + ",")
    out.append("};")
    open(os.path.join(HERE, "boys_table.h"), "w").write("\n".join(out) + "\n")
    print("boys_table.h: %d grid points x %d orders; asymptotic error %.2e" % (ng, MTAB, float(worst)))


if __name__ == "__main__":
    main()
