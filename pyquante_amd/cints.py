"""Drop-in for PyQuante's C extension module `cints` (Src/cints.c:572-803
(contr_coulomb, coulomb_repulsion), 552-557 (ijkl2intindex), 340-346
(Fgamma)).

Same function names and argument formats.  contr_coulomb and
coulomb_repulsion evaluate the Taketa-Huzinaga-O-ohata ERI of cints.c:80-153
on the GPU: each primitive quartet's integrand is expanded in u = t^2 and its
coefficients are contracted with the reference's EPS = 3e-7 Boys values, which
reproduces cints' numbers (they differ from crys' by that Boys error, ~1e-7).
Set `accurate_boys = True` for full-precision Boys values instead.
"""
from ._compat import THO
from ._compat import contr_coulomb as _contr
from ._compat import coulomb_repulsion as _prim
from ._compat import fgamma as _fgamma
from ._compat import intindex as ijkl2intindex  # noqa: F401

accurate_boys = False


def contr_coulomb(*args):
    return _contr(*args, method=THO, compat_nr=0 if accurate_boys else 1)


def coulomb_repulsion(*args):
    return _prim(*args, method=THO, compat_nr=0 if accurate_boys else 1)


def Fgamma(m, x):
    return _fgamma(m, x, compat_nr=1)
