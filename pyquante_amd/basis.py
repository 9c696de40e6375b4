"""Basis sets: contracted Cartesian Gaussian functions and shells.

Mirrors PyQuante/Basis/basis.py:11-50 (BasisSet), PyQuante/CGBF.py:27-75
(CGBF) and the primitive normalisation of Src/lib/primitive-gto.c:76-83:
  - functions are ordered atom -> shell (basis-file order) -> Cartesian
    component in sym2powerlist order (Ints.py:23-29), D = xx,xy,xz,yy,yz,zz;
  - each primitive carries N = sqrt(2^(2L+1.5) a^(L+1.5) /
    ((2l-1)!!(2m-1)!!(2n-1)!! pi^1.5)) for its own (l,m,n);
  - contraction coefficients are the raw basis-file values and the
    contracted function is NOT renormalised (norm = 1.0, CGBF.py:34).
"""
import json
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))

sym2powerlist = {
    "S": [(0, 0, 0)],
    "P": [(1, 0, 0), (0, 1, 0), (0, 0, 1)],
    "D": [(2, 0, 0), (1, 1, 0), (1, 0, 1), (0, 2, 0), (0, 1, 1), (0, 0, 2)],
}
sym2l = {"S": 0, "P": 1, "D": 2}

basis_map = {"sto3g": "sto3g", "sto-3g": "sto3g", "6-31g**": "p631ss",
             "6-31g(d,p)": "p631ss", "p631ss": "p631ss"}

_cache = {}


def get_basis_data(name):
    """Basis data by name, as {Z: [(sym, [(exp, coef), ...]), ...]} (Basis/Tools.py:47-51)."""
    key = basis_map.get(name.lower())
    if key is None:
        raise ValueError("Can't import basis set %s" % name.lower())
    if key not in _cache:
        with open(os.path.join(_HERE, "data", key + ".json")) as f:
            raw = json.load(f)
        _cache[key] = {int(z): [(sym, [tuple(p) for p in prims]) for sym, prims in shells]
                       for z, shells in raw.items()}
    return _cache[key]


def _fact2(n):
    return 1 if n <= 1 else n * _fact2(n - 2)


def prim_norm(alpha, powers):
    l, m, n = powers
    L = l + m + n
    return math.sqrt(2 ** (2 * L + 1.5) * alpha ** (L + 1.5) /
                     _fact2(2 * l - 1) / _fact2(2 * m - 1) / _fact2(2 * n - 1) / math.pi ** 1.5)


class CGBF(object):
    """Contracted Gaussian basis function with the attributes CGBF.coulomb reads
    (pexps, pcoefs, pnorms, origin, powers, norm; CGBF.py:27-75)."""

    def __init__(self, origin, powers=(0, 0, 0), atid=0):
        self.origin = tuple(float(v) for v in origin)
        self.powers = tuple(int(p) for p in powers)
        self.norm = 1.0
        self.pexps = []
        self.pcoefs = []
        self.pnorms = []
        self.ang_mom = sum(self.powers)
        self.atid = atid

    def add_primitive(self, exponent, coefficient):
        self.pexps.append(float(exponent))
        self.pcoefs.append(float(coefficient))
        self.pnorms.append(prim_norm(float(exponent), self.powers))

    def nprims(self):
        return len(self.pexps)


class Shell(object):
    """A shell: L, exponents, raw coefficients, centre, atom and first function."""

    def __init__(self, sym, origin, atid, first_bf, prims):
        self.sym = sym
        self.L = sym2l[sym]
        self.origin = tuple(float(v) for v in origin)
        self.atid = atid
        self.first_bf = first_bf
        self.exps = [float(e) for e, _ in prims]
        self.coefs = [float(c) for _, c in prims]

    @property
    def nfunc(self):
        return len(sym2powerlist[self.sym])


class BasisSet(object):
    """BasisSet(atoms, basis_data=None) with .bfs and .shells (basis.py:11-50).
    basis_data may be a name ('sto-3g', '6-31G**') or a data dict; the
    default is 6-31G**, as in the reference."""

    def __init__(self, atoms, basis_data=None, **opts):
        if not basis_data:
            basis_data = get_basis_data("6-31g**")
        elif isinstance(basis_data, str):
            basis_data = get_basis_data(basis_data)
        self.bfs = []
        self.shells = []
        self.atoms = list(atoms)
        for atom in self.atoms:
            for sym, prims in basis_data[atom.atno]:
                if sym not in sym2powerlist:
                    raise ValueError("only S, P and D shells are supported (got %s)" % sym)
                sh = Shell(sym, atom.pos(), atom.atid, len(self.bfs), prims)
                sh.basis, sh.index = self, len(self.shells)
                for power in sym2powerlist[sym]:
                    bf = CGBF(atom.pos(), power, atom.atid)
                    for e, c in prims:
                        bf.add_primitive(e, c)
                    self.bfs.append(bf)
                self.shells.append(sh)

    def __len__(self):
        return len(self.bfs)

    def __iter__(self):
        return iter(self.bfs)

    def __getitem__(self, i):
        return self.bfs[i]


def getbasis(atoms, basis_data=None, **opts):
    """Ints.getbasis (Ints.py:35-43): returns the BasisSet."""
    return BasisSet(atoms, basis_data, **opts)
