"""Shared implementation of the drop-in cints / crys / chgp module functions.

Argument formats follow the reference's C wrappers (crys.c:1556-1711,
cints.c:552-803, chgp.c:701-855): contr_coulomb takes 20 arguments
(exps, coefs, norms, (x,y,z), (l,m,n)) x 4 and coulomb_repulsion takes
((x,y,z), norm, (l,m,n), alpha) x 4.  Values are computed on the GPU.
The reference returns NULL without an exception for malformed input and
more than 40 primitives in total (crys.c:1579-1633); that surfaces in
Python as SystemError, which is raised here in the same cases.
"""
import ctypes

import numpy as np

from . import _abi
from ._abi import check, lib, pd, pi

MAX_TOTAL_PRIMS = 40


def _seq(x, n=None):
    try:
        v = [float(t) for t in x]
    except TypeError:
        raise SystemError("error return without exception set")
    if n is not None and len(v) != n:
        raise SystemError("error return without exception set")
    return v


THO, RYS, HGP = 0, 1, 2


def contr_coulomb_batch(quartets, method=RYS, compat_nr=1, mext=None):
    """quartets: list of 4-tuples of (exps, coefs, norms, origin, powers).

    method RYS evaluates with the Rys engine (full precision); THO contracts
    the integrand's coefficients in u with Boys values, and with compat_nr = 1
    (the EPS = 3e-7 Boys function of cints.c) reproduces cints.c's values; HGP
    runs the Head-Gordon-Pople kernel, whose compat_nr = 1 reproduces chgp.c's
    values.  mext: per-quartet auxiliary order for primitive VRR leaves (HGP
    only)."""
    _abi.require_gpu()
    nprim, poff, exps, coefs, norms, org, pw = [], [], [], [], [], [], []
    for q in quartets:
        if len(q) != 4:
            raise SystemError("error return without exception set")
        tot = 0
        for (e, c, nrm, o, p) in q:
            e, c, nrm = _seq(e), _seq(c), _seq(nrm)
            if not (len(e) == len(c) == len(nrm)) or len(e) == 0:
                raise SystemError("error return without exception set")
            tot += len(e)
            poff.append(len(exps))
            nprim.append(len(e))
            exps += e
            coefs += c
            norms += nrm
            org += _seq(o, 3)
            pw += [int(t) for t in p]
        if tot > MAX_TOTAL_PRIMS:
            raise SystemError("error return without exception set")
    n = len(quartets)
    out = np.zeros(n)
    a = [np.array(x, dtype=t) for x, t in ((nprim, np.int32), (poff, np.int32), (exps, np.float64),
                                           (coefs, np.float64), (norms, np.float64), (org, np.float64),
                                           (pw, np.int32))]
    try:
        if method == HGP:
            m = None if mext is None else np.ascontiguousarray(np.asarray(mext, dtype=np.int32).reshape(n))
            check(lib.pq_hgp_contr_coulomb_batch(n, pi(a[0]), pi(a[1]), len(exps), pd(a[2]), pd(a[3]), pd(a[4]),
                                                 pd(a[5]), pi(a[6]), None if m is None else pi(m), int(compat_nr),
                                                 pd(out)))
        elif method == THO:
            check(lib.pq_tho_contr_coulomb_batch(n, pi(a[0]), pi(a[1]), len(exps), pd(a[2]), pd(a[3]), pd(a[4]),
                                                 pd(a[5]), pi(a[6]), int(compat_nr), pd(out)))
        else:
            check(lib.pq_contr_coulomb_batch(n, pi(a[0]), pi(a[1]), len(exps), pd(a[2]), pd(a[3]), pd(a[4]),
                                             pd(a[5]), pi(a[6]), pd(out)))
    except _abi.PQError as e:
        if e.code in (-2, -3):
            raise ValueError(str(e))
        raise
    return out


def contr_coulomb(*args, method=RYS, compat_nr=1):
    if len(args) != 20:
        raise SystemError("error return without exception set")
    q = tuple(tuple(args[5 * f:5 * f + 5]) for f in range(4))
    return float(contr_coulomb_batch([q], method, compat_nr)[0])


def coulomb_repulsion(*args, method=RYS, compat_nr=1):
    if len(args) != 16:
        raise SystemError("error return without exception set")
    q = []
    for f in range(4):
        o, nrm, p, a = args[4 * f:4 * f + 4]
        q.append(([float(a)], [1.0], [float(nrm)], o, p))
    return float(contr_coulomb_batch([tuple(q)], method, compat_nr)[0])


def hgp_vrr(*args):
    """chgp.vrr (chgp.c:857-890): (A, norma, (la,ma,na), alphaa, B, normb,
    alphab, C, normc, (lc,mc,nc), alphac, D, normd, alphad, m) -> [a0|c0]^(m)."""
    if len(args) != 15:
        raise SystemError("error return without exception set")
    A, na, pa, aa, B, nb, ab, C, nc, pc, ac, D, nd, ad, m = args
    q = (([float(aa)], [1.0], [float(na)], A, pa), ([float(ab)], [1.0], [float(nb)], B, (0, 0, 0)),
         ([float(ac)], [1.0], [float(nc)], C, pc), ([float(ad)], [1.0], [float(nd)], D, (0, 0, 0)))
    return float(contr_coulomb_batch([q], HGP, 1, [int(m)])[0])


def intindex(i, j, k, l):
    out = ctypes.c_uint64()
    check(lib.pq_ijkl2intindex(int(i), int(j), int(k), int(l), ctypes.byref(out)))
    return int(out.value)


def fgamma(m, x, compat_nr=1):
    _abi.require_gpu()
    mm = np.array([float(m)])
    xx = np.array([float(x)])
    out = np.zeros(1)
    check(lib.pq_boys_batch(1, pd(mm), pd(xx), int(compat_nr), pd(out)))
    return float(out[0])
