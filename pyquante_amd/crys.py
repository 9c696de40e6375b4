"""Drop-in for PyQuante's C extension module `crys` (Src/crys.c:1556-1711 (contr_coulomb, coulomb_repulsion)).

Same function names and argument formats; the integrals are evaluated by the
GPU Rys engine, so every method returns the same (full double precision)
values.  Fgamma keeps the reference's Numerical-Recipes stopping rule
(EPS = 3e-7) so that it reproduces the reference's values.
"""
from ._compat import contr_coulomb, coulomb_repulsion, fgamma as _fgamma, intindex as ijkl2intindex  # noqa: F401


def Fgamma(m, x):
    return _fgamma(m, x, compat_nr=1)
