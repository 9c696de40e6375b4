"""Many independent molecules through one engine.

All molecules' shells go into one engine (pq_engine_create_batch): each
class kernel is one launch over the flat (molecule, class, tile) work list of
the whole batch, and D, J, K, S, T, V and W are packed per molecule (molecule
m's n_m x n_m block at offset moff[m]).  The SCF runs for all molecules at
once: the Fock build is one HIP-graph replay over the packed density, the
matrix algebra runs on the blocks padded to the largest molecule (padding
rows carry an identity overlap and a large diagonal, so their orbitals sort
last and are never occupied), and the eigensolves run batched per group of
equal-size molecules on their unpadded blocks.
"""
import ctypes

import numpy as np
import torch

from . import _abi
from ._abi import check, lib
from .basis import BasisSet
from .engine import Engine
from .hartree_fock import mkdens  # noqa: F401  (same density convention)


class BatchEngine(Engine):
    """BatchEngine(mols, basis_data="6-31g**"): one engine over a list of
    Molecule objects."""

    def __init__(self, mols, basis_data="6-31g**"):
        atoms, atom_mol, shells, shell_atoms = [], [], [], []
        for m, mol in enumerate(mols):
            base = len(atoms)
            mol_atoms = list(mol)
            bs = BasisSet(mol_atoms, basis_data)
            index = {id(a): k for k, a in enumerate(mol_atoms)}
            for sh in bs.shells:
                shells.append((sh.L, sh.exps, sh.coefs, sh.atid, sh.origin))
                # the shell's atom: the atom of this molecule at its centre
                d = [np.sum((np.array(a.pos()) - np.array(sh.origin)) ** 2) for a in mol_atoms]
                shell_atoms.append(base + int(np.argmin(d)))
            atoms += mol_atoms
            atom_mol += [m] * len(mol_atoms)
            del index
        super().__init__(None, atoms, atom_mol=atom_mol, shells=shells, shell_atoms=shell_atoms)
        self.mols = list(mols)
        nm = self.nmol
        nbf = (ctypes.c_int * nm)()
        moff = (ctypes.c_int64 * nm)()
        at0 = (ctypes.c_int * nm)()
        nat = (ctypes.c_int * nm)()
        n_, ms_ = ctypes.c_int(), ctypes.c_int64()
        check(lib.pq_engine_layout(self.handle, ctypes.byref(n_), ctypes.byref(ms_), nbf, moff, at0, nat))
        self.mol_nbf = np.array(nbf[:], dtype=np.int64)
        self.mol_moff = np.array(moff[:], dtype=np.int64)
        self.mol_atom0 = np.array(at0[:], dtype=np.int64)
        self.mol_natom = np.array(nat[:], dtype=np.int64)
        self.nmax = int(self.mol_nbf.max())
        # packed element -> position in the padded (nmol, nmax, nmax) batch
        idx = []
        for m in range(nm):
            n = int(self.mol_nbf[m])
            r = np.arange(n)
            idx.append(m * self.nmax * self.nmax + r[:, None] * self.nmax + r[None, :])
        self.pad_index = torch.as_tensor(np.concatenate([i.reshape(-1) for i in idx]), device=self.device)
        mask = np.zeros((nm, self.nmax), dtype=bool)
        for m in range(nm):
            mask[m, :self.mol_nbf[m]] = True
        self.valid = torch.as_tensor(mask, device=self.device)

    # packed <-> padded batch
    def pad(self, packed, fill_diag=0.0):
        out = torch.zeros((self.nmol * self.nmax * self.nmax,), dtype=packed.dtype, device=packed.device)
        out[self.pad_index] = packed
        out = out.view(self.nmol, self.nmax, self.nmax)
        if fill_diag:
            d = torch.diagonal(out, dim1=1, dim2=2)
            d[~self.valid] = fill_diag
        return out

    def pack(self, padded):
        return padded.reshape(-1)[self.pad_index].contiguous()

    def blocks(self, packed):
        """Per-molecule n_m x n_m views of a packed matrix."""
        return [packed[o:o + n * n].view(n, n) for o, n in zip(self.mol_moff, self.mol_nbf)]


_CHUNK = 64  # molecules per batched linear-algebra call (library workspace limits)


def _chunked(fn, *xs):
    """fn over slices of at most _CHUNK molecules, results concatenated (a
    tuple of outputs is concatenated element-wise)."""
    outs = [fn(*(x[i:i + _CHUNK] for x in xs)) for i in range(0, xs[0].shape[0], _CHUNK)]
    if isinstance(outs[0], tuple):
        return tuple(torch.cat(o, 0) for o in zip(*outs))
    return torch.cat(outs, 0)


class _DiisState(object):
    """History of one batched DIIS: Fock matrices, error vectors and the Gram
    matrix of the errors, one row added per iteration (the stored errors are
    not multiplied again)."""

    def __init__(self):
        self.Fs, self.Es, self.G = [], [], None


def _batched_diis_fock(st, F, err, nkeep=8):
    """Pulay DIIS per molecule on padded batches: returns the extrapolated
    Fock matrices (nmol, nmax, nmax)."""
    nm = F.shape[0]
    if st.G is None:
        st.G = torch.zeros((nm, nkeep, nkeep), dtype=F.dtype, device=F.device)
    if len(st.Fs) == nkeep:  # drop the oldest entry and its Gram row and column
        st.Fs.pop(0)
        st.Es.pop(0)
        st.G[:, :-1, :-1] = st.G[:, 1:, 1:].clone()
    e = err.reshape(nm, -1)
    st.Fs.append(F)
    st.Es.append(e)
    k = len(st.Fs)
    for j in range(k):
        g = (st.Es[j] * e).sum(1)
        st.G[:, k - 1, j] = g
        st.G[:, j, k - 1] = g
    if k < 2:
        return F
    # the (k+1) x (k+1) systems are solved on the host (one small copy; the
    # batched GPU solver's first call alone took 0.28 s)
    B = -np.ones((nm, k + 1, k + 1))
    B[:, :k, :k] = st.G[:, :k, :k].cpu().numpy()
    B[:, k, k] = 0.0
    rhs = np.zeros((nm, k + 1, 1))
    rhs[:, k, 0] = -1.0
    # scale each molecule's system for conditioning
    scale = np.maximum(np.abs(np.diagonal(B[:, :k, :k], axis1=1, axis2=2)).max(1), 1e-300)
    B[:, :k, :k] /= scale[:, None, None]
    c = torch.as_tensor(np.linalg.solve(B, rhs)[:, :k, 0], dtype=F.dtype, device=F.device)
    out = c[:, 0, None, None] * st.Fs[0]
    for j in range(1, k):
        out += c[:, j, None, None] * st.Fs[j]
    return out


def batch_rhf(be, ConvCriteria=1e-8, MaxIter=60, precision=_abi.PQ_PREC_FP64, graph=True, diis=True,
              pad_energy=1e3):
    """Closed-shell RHF for every molecule of a BatchEngine at once
    (hartree_fock.rhf semantics per molecule, DIIS on by default).  Returns
    (energies[nmol], orbe (nmol, nmax), C (nmol, nmax, nmax), D packed,
    W packed, iterations)."""
    S, T, V = be.get1ints()
    Sp = be.pad(S, fill_diag=1.0)
    hp = be.pad(T + V, fill_diag=pad_energy)
    L = _chunked(torch.linalg.cholesky, Sp)
    eye = torch.eye(be.nmax, dtype=Sp.dtype, device=Sp.device).expand_as(Sp).contiguous()
    X = _chunked(lambda a, b: torch.linalg.solve_triangular(a, b, upper=False), L, eye).transpose(1, 2)
    nocc = torch.as_tensor([m.get_closedopen()[0] for m in be.mols], device=be.device)
    occ = (torch.arange(be.nmax, device=be.device)[None, :] < nocc[:, None]).to(torch.float64)
    enuke = torch.as_tensor([m.get_enuke() for m in be.mols], dtype=torch.float64, device=be.device)

    # the eigensolves run per group of equal-size molecules on their n x n
    # blocks (the padding block of X^T F X is diagonal and decoupled: its
    # eigenvectors are unit vectors, its eigenvalues the padded diagonal,
    # which sorts after every real one)
    groups = []
    for n in sorted(set(int(x) for x in be.mol_nbf)):
        idx = torch.as_tensor(np.nonzero(be.mol_nbf == n)[0], device=be.device)
        groups.append((n, idx))
    eye_pad = torch.eye(be.nmax, dtype=Sp.dtype, device=Sp.device)

    def solve(Fp):
        A = X.transpose(1, 2) @ Fp @ X
        e = torch.diagonal(A, dim1=1, dim2=2).clone()
        V = eye_pad.expand_as(A).clone()
        for n, idx in groups:
            eg, vg = _chunked(torch.linalg.eigh, A[idx, :n, :n])
            e[idx, :n] = eg
            V[idx, :n, :n] = vg
        return e, X @ V

    orbe, C = solve(hp)
    eold = torch.zeros_like(enuke)
    fg = None
    diis_state = _DiisState()
    it = 0
    energy = eold
    for it in range(1, MaxIter + 1):
        Dp = (C * occ[:, None, :]) @ C.transpose(1, 2)
        D = be.pack(Dp)
        if graph:
            if fg is None:
                fg = be.capture_fock_jk(D, precision=precision)
            J, K = fg(D)
        else:
            J, K = be.fock_jk(D, precision=precision)
        Fp = hp + be.pad(2.0 * J - K)
        energy = (Dp * (hp + Fp)).sum((1, 2)) + enuke
        if diis:
            err = Fp @ Dp @ Sp - Sp @ Dp @ Fp
            Fp = _batched_diis_fock(diis_state, Fp, err)
        orbe, C = solve(Fp)
        if float((energy - eold).abs().max()) < ConvCriteria:
            break
        eold = energy
    Dp = (C * occ[:, None, :]) @ C.transpose(1, 2)
    Wp = (C * (occ * orbe)[:, None, :]) @ C.transpose(1, 2)
    return energy, orbe, C, be.pack(Dp), be.pack(Wp), it


def batch_forces(be, D, W):
    """dE/dR of every molecule: a list of (natom_m, 3) tensors."""
    g = be.gradient(D, W)
    return [g[a0:a0 + na] for a0, na in zip(be.mol_atom0, be.mol_natom)]


def seeded_water_batch(nmol=512, nmax=12, seed=0):
    """The batched workload: nmol seeded water clusters (H2O)_n, n uniform in
    1..nmax, each from the cluster generator with its own seed."""
    from .molecule import lattice_for, water_cluster
    rng = np.random.default_rng(seed)
    ns = rng.integers(1, nmax + 1, size=nmol)
    return [water_cluster(lattice_for(int(n)), seed=seed * 1000003 + i, nwater=int(n)) for i, n in enumerate(ns)]
