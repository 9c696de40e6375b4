"""ctypes binding of the CPU oracle (liboracle.so) and of the compiled
reference harness (_ref/libpqref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker.  The product package never
imports this module.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_D = ctypes.c_double
_I = ctypes.c_int
_PD = ctypes.POINTER(ctypes.c_double)
_PI = ctypes.POINTER(ctypes.c_int)

THO, RYS, HGP = 0, 1, 2


def _pd(a):
    return a.ctypes.data_as(_PD)


def _pi(a):
    return a.ctypes.data_as(_PI)


def load():
    # PQ_ORACLE_LIB selects another build of the same source (the sanitizer
    # build, liboracle_san.so)
    path = os.path.join(HERE, os.environ.get("PQ_ORACLE_LIB", "liboracle.so"))
    if not os.path.exists(path):
        raise RuntimeError("oracle not built: run make -C oracle")
    lib = ctypes.CDLL(path)
    lib.pqo_fgamma_nr.restype = _D
    lib.pqo_fgamma_nr.argtypes = [_D, _D]
    lib.pqo_boys.restype = _D
    lib.pqo_boys.argtypes = [_D, _D]
    lib.pqo_rys_roots.restype = _I
    lib.pqo_rys_roots.argtypes = [_I, _D, _PD, _PD]
    for name in ("pqo_rys_coulomb_repulsion",):
        f = getattr(lib, name)
        f.restype = _D
        f.argtypes = [_PD, _D, _PI, _D] * 4
    for name in ("pqo_tho_coulomb_repulsion", "pqo_hgp_coulomb_repulsion"):
        f = getattr(lib, name)
        f.restype = _D
        f.argtypes = [_PD, _D, _PI, _D] * 4 + [_I]
    lib.pqo_contr_coulomb.restype = _D
    lib.pqo_contr_coulomb.argtypes = [_I, _I, _PI, _PD, _PD, _PD, _PD, _PI]
    lib.pqo_ijkl2intindex.restype = ctypes.c_uint64
    lib.pqo_ijkl2intindex.argtypes = [ctypes.c_int64] * 4
    lib.pqo_get2ints.restype = _I
    lib.pqo_get2ints.argtypes = [_I, _I, _I, _PI, _PI, _PD, _PD, _PD, _PD, _PI, _PD]
    lib.pqo_getJK.restype = None
    lib.pqo_getJK.argtypes = [_I, _PD, _PD, _PD, _PD, _PD]
    lib.pqo_overlap.restype = _D
    lib.pqo_overlap.argtypes = [_D, _PI, _PD, _D, _PI, _PD]
    lib.pqo_kinetic.restype = _D
    lib.pqo_kinetic.argtypes = [_D, _PI, _PD, _D, _PI, _PD]
    lib.pqo_nuclear_attraction.restype = _D
    lib.pqo_nuclear_attraction.argtypes = [_PD, _D, _PI, _D, _PD, _D, _PI, _D, _PD, _I]
    lib.pqo_get1ints.restype = None
    lib.pqo_get1ints.argtypes = [_I, _I, _PI, _PI, _PD, _PD, _PD, _PD, _PI, _I, _PD, _PD, _PD, _PD]
    lib.pqo_prim_norm.restype = _D
    lib.pqo_prim_norm.argtypes = [_D, _I, _I, _I]
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = load()
    return _lib


def boys(m, x):
    return lib().pqo_boys(float(m), float(x))


def fgamma_nr(m, x):
    return lib().pqo_fgamma_nr(float(m), float(x))


def rys_roots(n, X):
    u = np.zeros(n)
    w = np.zeros(n)
    rc = lib().pqo_rys_roots(n, float(X), _pd(u), _pd(w))
    if rc:
        raise ValueError("bad root count %d" % n)
    return u, w


def ijkl2intindex(i, j, k, l):
    return int(lib().pqo_ijkl2intindex(i, j, k, l))


class FlatBasis(object):
    """Contracted functions flattened into arrays (the oracle's input)."""

    def __init__(self, bfs):
        self.nbf = len(bfs)
        nprim = [len(b.pexps) for b in bfs]
        self.nprim = np.array(nprim, dtype=np.int32)
        self.poff = np.concatenate([[0], np.cumsum(nprim)[:-1]]).astype(np.int32)
        self.exps = np.array([e for b in bfs for e in b.pexps], dtype=np.float64)
        self.coefs = np.array([c for b in bfs for c in b.pcoefs], dtype=np.float64)
        self.norms = np.array([n for b in bfs for n in b.pnorms], dtype=np.float64)
        self.org = np.array([b.origin for b in bfs], dtype=np.float64).reshape(-1)
        self.pw = np.array([b.powers for b in bfs], dtype=np.int32).reshape(-1)

    def args(self):
        return (self.nbf, _pi(self.nprim), _pi(self.poff), _pd(self.exps), _pd(self.coefs),
                _pd(self.norms), _pd(self.org), _pi(self.pw))


def get2ints(bfs, method=RYS, compat_nr=0):
    fb = FlatBasis(bfs)
    n = fb.nbf
    npair = n * (n + 1) // 2
    out = np.zeros(npair * (npair + 1) // 2)
    lib().pqo_get2ints(method, compat_nr, *fb.args(), _pd(out))
    return out


def getJK(ints, D):
    D = np.ascontiguousarray(D, dtype=np.float64)
    n = D.shape[0]
    J = np.zeros((n, n))
    K = np.zeros((n, n))
    G = np.zeros((n, n))
    lib().pqo_getJK(n, _pd(np.ascontiguousarray(ints)), _pd(D), _pd(J), _pd(K), _pd(G))
    return J, K, G


def get1ints(bfs, atoms, compat_nr=0):
    fb = FlatBasis(bfs)
    n = fb.nbf
    xyz = np.array([a.pos() for a in atoms], dtype=np.float64).reshape(-1)
    z = np.array([a.atno for a in atoms], dtype=np.float64)
    S = np.zeros((n, n))
    h = np.zeros((n, n))
    lib().pqo_get1ints(compat_nr, *fb.args(), len(atoms), _pd(xyz), _pd(z), _pd(S), _pd(h))
    return S, h


def contr_coulomb(a, b, c, d, method=RYS, compat_nr=0):
    """Contracted ERI of four function objects exposing pexps/pcoefs/pnorms/
    origin/powers (the CGBF attributes the reference's CGBF.coulomb reads)."""
    fs = (a, b, c, d)
    nprim = np.array([len(f.pexps) for f in fs], dtype=np.int32)
    exps = np.array([e for f in fs for e in f.pexps])
    coefs = np.array([e for f in fs for e in f.pcoefs])
    norms = np.array([e for f in fs for e in f.pnorms])
    org = np.array([f.origin for f in fs], dtype=np.float64).reshape(-1)
    pw = np.array([f.powers for f in fs], dtype=np.int32).reshape(-1)
    return lib().pqo_contr_coulomb(method, compat_nr, _pi(nprim), _pd(exps), _pd(coefs),
                                   _pd(norms), _pd(org), _pi(pw))


def prim_eri(method, A, na, pa, aa, B, nb, pb, ab, C, nc, pc, ac, D, nd, pd, ad, compat_nr=0):
    args = []
    for (X, n, p, a) in ((A, na, pa, aa), (B, nb, pb, ab), (C, nc, pc, ac), (D, nd, pd, ad)):
        args += [_pd(np.asarray(X, dtype=np.float64)), float(n),
                 _pi(np.asarray(p, dtype=np.int32)), float(a)]
    keep = args  # noqa: F841  (arrays must outlive the call)
    if method == RYS:
        return lib().pqo_rys_coulomb_repulsion(*args)
    if method == THO:
        return lib().pqo_tho_coulomb_repulsion(*(args + [compat_nr]))
    return lib().pqo_hgp_coulomb_repulsion(*(args + [compat_nr]))
