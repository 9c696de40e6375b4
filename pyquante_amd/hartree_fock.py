"""Closed-shell Hartree-Fock driver on the GPU engine.

Mirrors PyQuante/hartree_fock.py:66-149 (rhf), LA2.py:47-131 (geigh with
Cholesky orthogonalisation, mkdens, trace2) and Convergence.py:73-136
(Pulay DIIS): the same options (ConvCriteria, MaxIter, DoAveraging, bfs,
basis_data, integrals, orbs), the same iteration (D = C_occ C_occ^T,
F = h + 2J - K, E = tr(D h) + tr(D F) + E_nuc, stop when |E - E_old| <
ConvCriteria with E_old starting at 0) and the same defaults (DIIS off).
All matrices are FP64 torch tensors on the GPU; the two-electron part is
the engine's direct Fock build.
"""
import logging

import numpy as np
import torch

from . import _abi
from .Ints import Integrals, getbasis
from .engine import Engine

logger = logging.getLogger("pyquante")


def _t(x, dev):
    if torch.is_tensor(x):
        return x.to(dev, torch.float64)
    return torch.as_tensor(np.asarray(x, dtype=np.float64), device=dev)


def chol_orth(S):
    """X = inv(L)^T with L L^T = S (LA2.py:118-121)."""
    L = torch.linalg.cholesky(S)
    return torch.linalg.inv(L).T


# Largest order solved by the host's LAPACK instead of the GPU eigensolver:
# small symmetric eigenproblems are latency-bound on the GPU (measured on
# MI355X: 4.3 ms at n = 200 against 2.6 ms on the host with the copies; the
# two meet near n = 400).
HOST_EIGH_MAX = 320


def geigh(H, S, X=None):
    """Generalised eigenproblem H C = S C e by Cholesky orthogonalisation
    (LA2.py:47-75).  Returns (eigenvalues, eigenvectors as columns)."""
    if X is None:
        X = chol_orth(S)
    A = X.T @ H @ X
    if A.shape[0] <= HOST_EIGH_MAX:
        w, v = np.linalg.eigh(A.cpu().numpy())
        e, v = torch.as_tensor(w, device=A.device), torch.as_tensor(v, device=A.device)
    else:
        e, v = torch.linalg.eigh(A)
    return e, X @ v


def mkdens(c, nstart, nstop):
    """D = C[:, nstart:nstop] C[:, nstart:nstop]^T (LA2.py:127-131)."""
    d = c[:, nstart:nstop]
    return d @ d.T


def trace2(A, B):
    return float(torch.sum(A * B))


def get_energy(h, F, D, enuke=0.0):
    """E = tr(D h) + tr(D F) + E_nuc (hartree_fock.py:33-38)."""
    return trace2(D, h) + trace2(D, F) + enuke


class DIIS(object):
    """Pulay DIIS on the commutator error FDS - SDF (Convergence.py:73-136),
    every stored error in the system as the reference keeps them.  The
    errors and Fock matrices live in growing device buffers; the Gram matrix
    of the errors gains one row per call (one matrix-vector product), and
    the small (k+1) x (k+1) system is solved on the host, as the reference's
    LAPACK solve is."""

    def __init__(self, S):
        self.S = S
        self.maxerr = None
        self.k = 0
        self._E = self._F = None
        self._G = np.zeros((0, 0))

    def _reserve(self, err, F):
        cap = 0 if self._E is None else self._E.shape[0]
        if self.k < cap:
            return
        ncap = max(8, 2 * cap)
        E = torch.empty((ncap, err.numel()), dtype=err.dtype, device=err.device)
        Fb = torch.empty((ncap,) + tuple(F.shape), dtype=F.dtype, device=F.device)
        if cap:
            E[:cap] = self._E
            Fb[:cap] = self._F
        self._E, self._F = E, Fb

    def getF(self, F, D):
        err = (F @ D @ self.S - self.S @ D @ F).reshape(-1)
        self._reserve(err, F)
        k = self.k
        self._E[k] = err
        self._F[k] = F
        nit = k + 1
        self.k = nit
        # new Gram row and the largest error element, one copy to the host
        row = torch.cat([self._E[:nit] @ err, err.abs().max().reshape(1)]).cpu().numpy()
        self.maxerr = float(row[-1])
        G = np.zeros((nit, nit))
        G[:k, :k] = self._G
        G[k, :] = row[:nit]
        G[:, k] = row[:nit]
        self._G = G
        a = np.zeros((nit + 1, nit + 1))
        a[:nit, :nit] = G
        a[nit, :nit] = -1.0
        a[:nit, nit] = -1.0
        b = np.zeros(nit + 1)
        b[nit] = -1.0
        try:
            c = np.linalg.solve(a, b)
        except np.linalg.LinAlgError:
            return F
        if not np.all(np.isfinite(c)):
            return F
        cd = torch.as_tensor(c[:nit], dtype=F.dtype, device=F.device)
        return torch.tensordot(cd, self._F[:nit], dims=1)


def rhf(atoms, **opts):
    """rhf(atoms, **opts) -> (energy, orbe, orbs) (hartree_fock.py:66-149).

    Extra options: method "rys" (direct J/K) or "hgp" (stored HGP integrals,
    as with the reference's chgp backend); precision PQ_PREC_MIXED (direct
    Rys with FP32 quartets below fp32_cut); reference_boys (default
    Ints.reference_boys: V with the reference's EPS = 3e-7 Boys function, as
    its getints computes it)."""
    from . import Ints as _Ints
    ConvCriteria = opts.get("ConvCriteria", 1e-4)
    MaxIter = opts.get("MaxIter", 20)
    DoAveraging = opts.get("DoAveraging", False)
    precision = opts.get("precision", _abi.PQ_PREC_FP64)
    bfs = opts.get("bfs", None)
    if not bfs:
        bfs = getbasis(atoms, opts.get("basis_data", None))
    nclosed, nopen = atoms.get_closedopen()
    if nopen:
        raise ValueError("SCF currently only works for closed-shell systems")
    nocc = nclosed
    integrals = opts.get("integrals", None)
    engine = None
    if integrals:
        S, h, Ints = integrals
        dev = Ints.engine.device if isinstance(Ints, Integrals) else torch.device("cuda")
    else:
        engine = Engine(bfs, atoms)
        dev = engine.device
        Sd, Td, Vd = engine.get1ints(1 if opts.get("reference_boys", _Ints.reference_boys) else 0)
        S, h = Sd, Td + Vd
        Ints = Integrals(bfs, atoms, engine=engine, method=opts.get("method", None))
    S = _t(S, dev)
    h = _t(h, dev)
    X = chol_orth(S)
    orbs = opts.get("orbs", None)
    if orbs is None:
        orbe, orbs = geigh(h, S, X)
    else:
        orbs = _t(orbs, dev)
    enuke = atoms.get_enuke()
    eold = 0.0
    avg = DIIS(S) if DoAveraging else None
    energy = 0.0
    mixed = precision != _abi.PQ_PREC_FP64
    if mixed:
        # mixed-precision work lists, built once for the whole SCF
        Ints.engine.build(opts.get("schwarz_cut", 1e-12), opts.get("fp32_cut", 1e-4))
    for i in range(MaxIter):
        D = mkdens(orbs, 0, nocc)
        if not mixed:
            G = _Ints.get2JmK(Ints, D)
        else:
            J, K = Ints.engine.fock_jk(0.5 * (D + D.T), precision=precision)
            G = 2.0 * J - K
        F = h + G
        if avg is not None:
            F = avg.getF(F, D)
        orbe, orbs = geigh(F, S, X)
        energy = get_energy(h, F, D, enuke)
        logger.debug("%d %f", i, energy)
        if abs(energy - eold) < ConvCriteria:
            break
        eold = energy
    if mixed:
        # later FP64 calls through Ints.jk rebuild their own lists
        Ints._built_cut = None
    return energy, orbe.cpu().numpy(), orbs.cpu().numpy()


def uhf(atoms, **opts):
    """uhf(atoms, **opts) -> (energy, (orbea, orbeb), (orbsa, orbsb))
    (hartree_fock.py:150-249): unrestricted Hartree-Fock with the reference's
    density averaging (DoAveraging, averaging = 0.5).  Each iteration's four
    getJ/getK calls are one two-density pass over the integrals on the GPU
    (pq_fock_jk_multi).  The reference's finite-temperature occupations
    (ETemp) are not supported."""
    if opts.get("ETemp", False):
        raise NotImplementedError("ETemp occupations are not supported")
    ConvCriteria = opts.get("ConvCriteria", 1e-5)
    MaxIter = opts.get("MaxIter", 40)
    DoAveraging = opts.get("DoAveraging", True)
    averaging = opts.get("averaging", 0.5)
    bfs = opts.get("bfs", None)
    if not bfs:
        bfs = getbasis(atoms, opts.get("basis_data", None))
    nalpha, nbeta = atoms.get_alphabeta()
    from . import Ints as _Ints
    engine = Engine(bfs, atoms)
    engine.build(opts.get("schwarz_cut", 1e-12), 0.0)
    dev = engine.device
    Sd, Td, Vd = engine.get1ints(1 if opts.get("reference_boys", _Ints.reference_boys) else 0)
    S, h = Sd, Td + Vd
    X = chol_orth(S)
    orbs = opts.get("orbs", None)
    if orbs is not None:
        orbsa, orbsb = _t(orbs[0], dev), _t(orbs[1], dev)
        orbea = orbeb = None
    else:
        orbe, orbs = geigh(h, S, X)
        orbea = orbeb = orbe
        orbsa = orbsb = orbs
    enuke = atoms.get_enuke()
    eold = 0.0
    energy = 0.0
    Da0 = Db0 = None
    for i in range(MaxIter):
        Da = mkdens(orbsa, 0, nalpha)
        Db = mkdens(orbsb, 0, nbeta)
        if DoAveraging:
            if i:
                Da = averaging * Da + (1 - averaging) * Da0
                Db = averaging * Db + (1 - averaging) * Db0
            Da0, Db0 = Da, Db
        (Ja, Ka), (Jb, Kb) = engine.fock_jk_multi([0.5 * (Da + Da.T), 0.5 * (Db + Db.T)])
        Fa = h + Ja + Jb - Ka
        Fb = h + Ja + Jb - Kb
        orbea, orbsa = geigh(Fa, S, X)
        orbeb, orbsb = geigh(Fb, S, X)
        energy = (get_energy(h, Fa, Da) + get_energy(h, Fb, Db)) / 2 + enuke
        logger.debug("%d %f", i, energy)
        if abs(energy - eold) < ConvCriteria:
            break
        eold = energy
    return energy, (orbea.cpu().numpy(), orbeb.cpu().numpy()), (orbsa.cpu().numpy(), orbsb.cpu().numpy())

