"""numpy restatement of PyQuante's closed-shell SCF (hartree_fock.py:66-149,
LA2.py:47-131) over the oracle's integrals.  TEST INFRASTRUCTURE ONLY."""
import numpy as np

from . import oracle as O


def geigh(H, S):
    X = np.linalg.inv(np.linalg.cholesky(S)).T
    e, v = np.linalg.eigh(X.T @ H @ X)
    return e, X @ v


def rhf(bfs, mol, ConvCriteria=1e-8, MaxIter=50, ints=None, S=None, h=None, compat_nr=1):
    """compat_nr = 1 (default): V with the reference's EPS = 3e-7 Boys
    function, as its getints computes it (Ints.py:73-87 over cints.c:241-275);
    0: V with the accurate Boys function (the V of the gradient kernels)."""
    if S is None:
        S, h = O.get1ints(bfs, mol.atoms, compat_nr=compat_nr)
    if ints is None:
        ints = O.get2ints(bfs)
    nocc = mol.get_closedopen()[0]
    orbe, orbs = geigh(h, S)
    enuke = mol.get_enuke()
    eold = 0.0
    energy = 0.0
    for _ in range(MaxIter):
        d = orbs[:, :nocc]
        D = d @ d.T
        _, _, G = O.getJK(ints, D)
        F = h + G
        orbe, orbs = geigh(F, S)
        energy = float(np.sum(D * h) + np.sum(D * F) + enuke)
        if abs(energy - eold) < ConvCriteria:
            break
        eold = energy
    return energy, orbe, orbs, D
