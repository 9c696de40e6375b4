"""Drop-in for the reference's libint module `clibint`
(Src/PyQuante/clibint.pyx, Src/lib/clibint.h) and the get2ints of its
libint branch (PyQuante/Ints.py:125-143).

shell_compute_eri(shell1, shell2, shell3, shell4, Ints) computes one shell
quartet and stores every integral at Ints[ijkl2intindex(i, j, k, l)] of the
basis-function indices, as the reference does; the block itself comes from
the GPU engine of the shells' basis set (pq_shell_quartet_blocks, libint's
row-major canonical-Cartesian layout).  The shells are the `.shells` of a
pyquante_amd BasisSet.
"""
import numpy as np


def _engine(basis):
    eng = getattr(basis, "_clibint_engine", None)
    if eng is None:
        from .engine import Engine
        eng = Engine(basis, basis.atoms)
        basis._clibint_engine = eng
    return eng


def _pair(p, q):
    hi, lo = np.maximum(p, q), np.minimum(p, q)
    return hi * (hi + 1) // 2 + lo


def _store(Ints, shells, block):
    r = [np.arange(sh.first_bf, sh.first_bf + sh.nfunc, dtype=np.int64) for sh in shells]
    ij = _pair(r[0][:, None], r[1][None, :])
    kl = _pair(r[2][:, None], r[3][None, :])
    idx = _pair(ij[:, :, None, None], kl[None, None, :, :])
    Ints[idx.reshape(-1)] = block.reshape(-1)


def shell_compute_eri(shell1, shell2, shell3, shell4, Ints):
    """clibint.pyx:27-60."""
    shells = (shell1, shell2, shell3, shell4)
    basis = shell1.basis
    if any(sh.basis is not basis for sh in shells):
        raise ValueError("the four shells must come from one basis set")
    block = _engine(basis).shell_quartets([[sh.index for sh in shells]])[0]
    _store(Ints, shells, block)


def get2ints(basis):
    """Ints.py:130-143 (libint branch): an n^4 array holding the unique ERIs
    at their ijkl2intindex positions, from the shell quartets the reference
    loops over (i >= j, k >= l, i + j >= k + l), evaluated in one batch."""
    n = len(basis.bfs)
    Ints = np.zeros(n ** 4, dtype=np.float64)
    sh = basis.shells
    quartets = [(i, j, k, l) for i in range(len(sh)) for j in range(i + 1) for k in range(len(sh))
                for l in range(k + 1) if i + j >= k + l]
    blocks = _engine(basis).shell_quartets(quartets)
    for q, block in zip(quartets, blocks):
        _store(Ints, [sh[t] for t in q], block)
    return Ints
