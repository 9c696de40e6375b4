"""Drop-in for PyQuante's C extension module `chgp` (Src/chgp.c:701-895).

Same function names and argument formats.  contr_coulomb (chgp.c:701-818)
and coulomb_repulsion (= hrr, chgp.c:820-855) run the GPU Head-Gordon-Pople
kernel, which evaluates the same horizontal and vertical recurrences with the
reference's Boys function (EPS = 3e-7, started at each leaf's own order), so
the values reproduce chgp's; vrr (chgp.c:857-890) returns one primitive leaf
[a0|c0]^(m).  Fgamma keeps the reference's Numerical-Recipes stopping rule.
Set `accurate_boys = True` for the full-precision Boys function instead.
"""
from ._compat import HGP, contr_coulomb_batch, fgamma as _fgamma, hgp_vrr as vrr  # noqa: F401
from ._compat import intindex as ijkl2intindex  # noqa: F401

accurate_boys = False


def _quartets_contr(args):
    if len(args) != 20:
        raise SystemError("error return without exception set")
    return tuple(tuple(args[5 * f:5 * f + 5]) for f in range(4))


def _quartets_prim(args):
    if len(args) != 16:
        raise SystemError("error return without exception set")
    q = []
    for f in range(4):
        o, nrm, p, a = args[4 * f:4 * f + 4]
        q.append(([float(a)], [1.0], [float(nrm)], o, p))
    return tuple(q)


def contr_coulomb(*args):
    return float(contr_coulomb_batch([_quartets_contr(args)], HGP, 0 if accurate_boys else 1)[0])


def coulomb_repulsion(*args):
    return float(contr_coulomb_batch([_quartets_prim(args)], HGP, 0 if accurate_boys else 1)[0])


hrr = coulomb_repulsion


def Fgamma(m, x):
    return _fgamma(m, x, compat_nr=1)
