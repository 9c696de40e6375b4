"""Hartree-Fock nuclear gradients and forces on the GPU engine.

Mirrors PyQuante/force.py:31-108 (hf_force / rhf_force / uhf_force /
fixedocc_uhf_force: forces -dE/dR_A stored on each atom).  For RHF every
term -- 2 tr(D dH), the two-electron derivative contracted with D on the
fly, -2 tr(Q dS) and dV_nn -- is computed by pq_forces, with
D = C_occ C_occ^T and Q (W) = C_occ diag(e_occ) C_occ^T
(Wavefunction.py:143-160); the unrestricted and fixed-occupation forces run
the same kernels on the spin densities (pq_forces_uhf).

RHFEnergy is a torch.autograd.Function: forward runs the SCF at the given
coordinates and returns E, keeping the engine and the converged D and W;
backward runs the gradient kernels on them and returns grad_output * dE/dR,
so E(R) composes with torch autograd and costs no gradient when none is
asked for.
"""
import numpy as np
import torch

from .basis import BasisSet
from .engine import Engine
from .hartree_fock import chol_orth, geigh, get_energy, mkdens


class Wavefunction(object):
    """Closed-shell wavefunction: orbitals (columns), energies, nclosed."""

    restricted = True
    unrestricted = False
    fixedocc = False

    def __init__(self, orbs, orbe, nclosed):
        self.orbs = orbs
        self.orbe = orbe
        self.nclosed = nclosed

    def mkdens(self):
        c = self.orbs[:, :self.nclosed]
        return c @ c.T

    def mkQmatrix(self):
        c = self.orbs[:, :self.nclosed]
        return (c * self.orbe[:self.nclosed]) @ c.T


class UHFWavefunction(object):
    """Unrestricted wavefunction: orbitals (orbsa, orbsb) (columns), energies
    (orbea, orbeb), nalpha and nbeta occupied orbitals; mkdens and mkQmatrix
    return the alpha and beta matrices (Wavefunction.py:143-160, unrestricted
    branch)."""

    restricted = False
    unrestricted = True
    fixedocc = False

    def __init__(self, orbs, orbe, nalpha, nbeta):
        self.orbs = orbs
        self.orbe = orbe
        self.nalpha = nalpha
        self.nbeta = nbeta

    def mkdens(self):
        ca = self.orbs[0][:, :self.nalpha]
        cb = self.orbs[1][:, :self.nbeta]
        return ca @ ca.T, cb @ cb.T

    def mkQmatrix(self):
        ca = self.orbs[0][:, :self.nalpha]
        cb = self.orbs[1][:, :self.nbeta]
        return (ca * self.orbe[0][:self.nalpha]) @ ca.T, (cb * self.orbe[1][:self.nbeta]) @ cb.T


def _vec(x, like):
    """x as a vector of like's kind (numpy array or device tensor)."""
    if isinstance(like, torch.Tensor):
        return torch.as_tensor(np.asarray(x, dtype=np.float64), device=like.device)
    return np.asarray(x, dtype=np.float64)


class FixedOccWavefunction(object):
    """Unrestricted wavefunction with fixed, possibly fractional, occupations
    (Wavefunction.py:161-199, the fixedocc branch): orbitals orbs_a, orbs_b
    (columns), energies orbe_a, orbe_b, occupation numbers occs_a, occs_b of
    the first len(occs) orbitals of each spin.  mk_auger_Qmatrix weights the
    beta orbitals with the alpha energies orbe_a, as the reference does
    (Wavefunction.py:197), so that forces match its fixedocc_uhf_force."""

    restricted = False
    unrestricted = False
    fixedocc = True

    def __init__(self, orbs_a, orbs_b, orbe_a, orbe_b, occs_a, occs_b):
        self.orbs_a, self.orbs_b = orbs_a, orbs_b
        self.orbe_a, self.orbe_b = orbe_a, orbe_b
        self.occs_a, self.occs_b = occs_a, occs_b

    def mk_auger_dens(self):
        fa, fb = _vec(self.occs_a, self.orbs_a), _vec(self.occs_b, self.orbs_b)
        ca, cb = self.orbs_a[:, :len(fa)], self.orbs_b[:, :len(fb)]
        return (ca * fa) @ ca.T, (cb * fb) @ cb.T

    def mk_auger_Qmatrix(self):
        fa, fb = _vec(self.occs_a, self.orbs_a), _vec(self.occs_b, self.orbs_b)
        ca, cb = self.orbs_a[:, :len(fa)], self.orbs_b[:, :len(fb)]
        ea = self.orbe_a
        return (ca * (ea[:len(fa)] * fa)) @ ca.T, (cb * (ea[:len(fb)] * fb)) @ cb.T


def scf(engine, mol, ConvCriteria=1e-10, MaxIter=100, DoAveraging=True, precision=0, graph=False):
    """RHF on an engine (hartree_fock.rhf semantics, DIIS on by default here
    for tight convergence).  Returns (E, orbe, orbs, D, W) as device tensors.
    graph=True replays each iteration's Fock build from one HIP graph
    captured on the first iteration (Engine.capture_fock_jk); the
    eigensolve stays outside the graph."""
    from .hartree_fock import DIIS
    S, T, V = engine.get1ints()
    h = T + V
    X = chol_orth(S)
    orbe, C = geigh(h, S, X)
    nocc = mol.get_closedopen()[0]
    enuke = mol.get_enuke()
    avg = DIIS(S) if DoAveraging else None
    eold = 0.0
    energy = 0.0
    fg = None
    for _ in range(MaxIter):
        D = mkdens(C, 0, nocc)
        if graph:
            if fg is None:
                fg = engine.capture_fock_jk(D, precision=precision)
            J, K = fg(D)
        else:
            J, K = engine.fock_jk(D, precision=precision)
        F = h + 2.0 * J - K
        if avg is not None:
            F = avg.getF(F, D)
        orbe, C = geigh(F, S, X)
        energy = get_energy(h, F, D, enuke)
        if abs(energy - eold) < ConvCriteria:
            break
        eold = energy
    D = mkdens(C, 0, nocc)
    W = (C[:, :nocc] * orbe[:nocc]) @ C[:, :nocc].T
    return energy, orbe, C, D, W


def rhf_gradient(mol, basis_data="6-31g**", ConvCriteria=1e-10, schwarz_cut=1e-12, engine=None):
    """(E, dE/dR) for a closed-shell molecule."""
    eng = engine if engine is not None else Engine(BasisSet(mol, basis_data), mol)
    eng.build(schwarz_cut, 0.0)
    E, orbe, C, D, W = scf(eng, mol, ConvCriteria=ConvCriteria)
    return E, eng.gradient(D, W)


def rhf_force(mol, wf, bset, engine=None):
    """force.py:48-70: sets atom.forces = -dE/dR_A for every atom."""
    eng = engine if engine is not None else Engine(bset, mol)
    if eng.cut is None:
        eng.build()
    D = torch.as_tensor(wf.mkdens(), dtype=torch.float64, device=eng.device)
    W = torch.as_tensor(wf.mkQmatrix(), dtype=torch.float64, device=eng.device)
    g = eng.gradient(D, W).cpu().numpy()
    for atom, gi in zip(mol.atoms, g):
        atom.forces = -gi
    return -g


def _spin_force(mol, dens, qmat, bset, engine):
    """-dE/dR from spin densities (Da, Db) and energy-weighted densities
    (Qa, Qb) by pq_forces_uhf; stored as atom.forces."""
    eng = engine if engine is not None else Engine(bset, mol)
    if eng.cut is None:
        eng.build()
    dev = eng.device
    Da, Db = (torch.as_tensor(x, dtype=torch.float64, device=dev) for x in dens)
    Qa, Qb = (torch.as_tensor(x, dtype=torch.float64, device=dev) for x in qmat)
    g = eng.gradient_uhf(Da, Db, Qa, Qb).cpu().numpy()
    for atom, gi in zip(mol.atoms, g):
        atom.forces = -gi
    return -g


def uhf_force(mol, wf, bset, engine=None):
    """force.py:72-89: sets atom.forces = -dE/dR_A for an unrestricted
    wavefunction (spin densities and energy-weighted densities of
    wf.mkdens / wf.mkQmatrix)."""
    return _spin_force(mol, wf.mkdens(), wf.mkQmatrix(), bset, engine)


def fixedocc_uhf_force(mol, wf, bset, engine=None):
    """force.py:91-108: the unrestricted forces with the fixed-occupation
    densities wf.mk_auger_dens() and wf.mk_auger_Qmatrix()."""
    return _spin_force(mol, wf.mk_auger_dens(), wf.mk_auger_Qmatrix(), bset, engine)


def uhf_gradient(mol, basis_data="6-31g**", ConvCriteria=1e-11, MaxIter=400, schwarz_cut=1e-12):
    """(E, dE/dR) for an unrestricted wavefunction of mol (its multiplicity),
    from hartree_fock.uhf."""
    from .hartree_fock import uhf
    bs = BasisSet(mol, basis_data)
    # full-precision Boys values in V: the one-electron derivative kernels
    # differentiate that V
    E, orbe, orbs = uhf(mol, bfs=bs.bfs, ConvCriteria=ConvCriteria, MaxIter=MaxIter, schwarz_cut=schwarz_cut,
                        reference_boys=False)
    na, nb = mol.get_alphabeta()
    wf = UHFWavefunction(orbs, orbe, na, nb)
    eng = Engine(bs, mol)
    eng.build(schwarz_cut, 0.0)
    Da, Db = (torch.as_tensor(x, device=eng.device) for x in wf.mkdens())
    Qa, Qb = (torch.as_tensor(x, device=eng.device) for x in wf.mkQmatrix())
    return E, eng.gradient_uhf(Da, Db, Qa, Qb)


def hf_force(mol, wf, bname):
    """force.py:31-46: the forces of the branch each of wf's flags
    (restricted, unrestricted, fixedocc) selects, in that order, each
    overwriting atom.forces; returns the last (None when no flag is set,
    as the reference then computes nothing)."""
    bset = BasisSet(mol, bname)
    f = None
    if getattr(wf, "restricted", False):
        f = rhf_force(mol, wf, bset)
    if getattr(wf, "unrestricted", False):
        f = uhf_force(mol, wf, bset)
    if getattr(wf, "fixedocc", False):
        f = fixedocc_uhf_force(mol, wf, bset)
    return f


class RHFEnergy(torch.autograd.Function):
    """E(R): forward = converged RHF energy at coordinates R (natom x 3, bohr),
    backward = dE/dR from the gradient kernels (pq_forces) on the converged
    D and energy-weighted W that forward keeps."""

    @staticmethod
    def forward(ctx, coords, mol, basis_data, conv):
        R = coords.detach().cpu().numpy()
        m = mol.copy_with_coords(R)
        eng = Engine(BasisSet(m, basis_data), m)
        eng.build(1e-12, 0.0)
        E, orbe, C, D, W = scf(eng, m, ConvCriteria=conv)
        ctx.engine = eng
        ctx.save_for_backward(D, W)
        return torch.tensor(E, dtype=coords.dtype, device=coords.device)

    @staticmethod
    def backward(ctx, grad_out):
        D, W = ctx.saved_tensors
        g = ctx.engine.gradient(D, W).to(grad_out.device, grad_out.dtype)
        ctx.engine = None
        return grad_out * g, None, None, None


def rhf_energy(coords, mol, basis_data="6-31g**", conv=1e-10):
    return RHFEnergy.apply(coords, mol, basis_data, conv)
