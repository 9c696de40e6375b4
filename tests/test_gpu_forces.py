"""GPU nuclear gradients against central finite differences of energies:
the oracle's CPU SCF (independent code) for H2O/STO-3G and the GPU's own
E(R) for 6-31G** water clusters; plus translational invariance."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from oracle import scf as oscf

pytestmark = pytest.mark.gpu


def _h2o_bent():
    from pyquante_amd.molecule import Molecule
    return Molecule("h2o", [(8, (0.0, 0.1, -0.05)), (1, (1.43, 1.05, 0.07)), (1, (-1.39, 1.12, -0.11))])


def test_gradient_vs_oracle_finite_difference_sto3g():
    from pyquante_amd.basis import BasisSet
    from pyquante_amd.force import rhf_gradient
    mol = _h2o_bent()
    E, g = rhf_gradient(mol, "sto-3g", ConvCriteria=1e-11, schwarz_cut=0.0)
    g = g.cpu().numpy()
    e_ref = oscf.rhf(BasisSet(mol, "sto-3g").bfs, mol, ConvCriteria=1e-11, MaxIter=200, compat_nr=0)[0]
    assert abs(E - e_ref) < 1e-9
    h = 1e-4
    X = mol.coords()
    for a in range(3):
        for x in range(3):
            Xp = X.copy(); Xp[a, x] += h
            Xm = X.copy(); Xm[a, x] -= h
            mp = mol.copy_with_coords(Xp)
            mm = mol.copy_with_coords(Xm)
            ep = oscf.rhf(BasisSet(mp, "sto-3g").bfs, mp, ConvCriteria=1e-12, MaxIter=200, compat_nr=0)[0]
            em = oscf.rhf(BasisSet(mm, "sto-3g").bfs, mm, ConvCriteria=1e-12, MaxIter=200, compat_nr=0)[0]
            fd = (ep - em) / (2 * h)
            assert abs(fd - g[a, x]) < 1e-6, (a, x, fd, g[a, x])
    assert np.max(np.abs(g.sum(0))) < 1e-8  # translational invariance


def test_gradient_vs_own_finite_difference_water_dimer_631ss():
    from pyquante_amd.force import rhf_gradient
    from pyquante_amd.molecule import water_cluster
    mol = water_cluster((2, 1, 1), seed=4)
    E, g = rhf_gradient(mol, "6-31g**", ConvCriteria=1e-11)
    g = g.cpu().numpy()
    assert np.max(np.abs(g.sum(0))) < 1e-7
    rng = np.random.default_rng(8)
    X = mol.coords()
    h = 1e-4
    for _ in range(4):
        a, x = int(rng.integers(len(mol))), int(rng.integers(3))
        Xp = X.copy(); Xp[a, x] += h
        Xm = X.copy(); Xm[a, x] -= h
        ep = rhf_gradient(mol.copy_with_coords(Xp), "6-31g**", ConvCriteria=1e-11)[0]
        em = rhf_gradient(mol.copy_with_coords(Xm), "6-31g**", ConvCriteria=1e-11)[0]
        fd = (ep - em) / (2 * h)
        assert abs(fd - g[a, x]) < 1e-6, (a, x, fd, g[a, x])


def test_autograd_function_and_hf_force():
    from pyquante_amd.basis import BasisSet
    from pyquante_amd.force import Wavefunction, hf_force, rhf_energy, rhf_gradient
    mol = _h2o_bent()
    R = torch.tensor(mol.coords(), dtype=torch.float64, requires_grad=True)
    E = rhf_energy(R, mol, "sto-3g")
    E.backward()
    E2, g = rhf_gradient(mol, "sto-3g")
    assert abs(float(E.detach()) - E2) < 1e-9
    assert torch.max(torch.abs(R.grad - g.cpu())).item() < 1e-9


def test_gradient_streams_agree():
    """The gradient kernels spread over the engine's streams give the gradient
    of the one-stream launch (to the reordering of the atomic sums)."""
    import numpy as np
    import torch
    from pyquante_amd.basis import BasisSet
    from pyquante_amd.engine import Engine
    from pyquante_amd.molecule import water_cluster
    mol = water_cluster((2, 1, 1), seed=6)
    eng = Engine(BasisSet(mol, "6-31g**"), mol)
    n = eng.nbf
    rng = np.random.default_rng(16)
    C = rng.standard_normal((n, 10)) * 0.1
    D = torch.as_tensor(C @ C.T, device=eng.device)
    W = torch.as_tensor(C @ np.diag(np.linspace(-1.0, -0.1, 10)) @ C.T, device=eng.device)
    eng.build(1e-12, 0.0)
    eng.set_streams(0)
    g0 = eng.gradient(D, W)
    eng.set_streams(4)
    g4 = eng.gradient(D, W)
    assert torch.max(torch.abs(g0 - g4)).item() < 1e-10
    # translational invariance holds for any D, W
    assert torch.max(torch.abs(g4.sum(dim=0))).item() < 1e-8


def _oh_radical():
    from pyquante_amd.molecule import Molecule
    return Molecule("oh", [(8, (0.0, 0.05, -0.02)), (1, (1.81, 0.12, 0.06))], multiplicity=2)


def test_uhf_gradient_closed_shell_equals_rhf():
    """force.py:72-89 on a closed shell: with equal spin densities the UHF
    gradient is the RHF one."""
    from pyquante_amd.force import rhf_gradient, uhf_gradient
    mol = _h2o_bent()
    e_r, g_r = rhf_gradient(mol, "sto-3g", ConvCriteria=1e-12)
    e_u, g_u = uhf_gradient(mol, "sto-3g", ConvCriteria=1e-12)
    assert abs(e_u - e_r) < 1e-9
    assert torch.max(torch.abs(g_u - g_r)).item() < 1e-7


def test_uhf_gradient_vs_own_finite_difference_oh():
    """UHF forces of the OH radical (doublet) against central finite
    differences (h = 1e-4 bohr) of the GPU UHF energy, every coordinate;
    translational invariance; hf_force dispatch on an unrestricted
    wavefunction."""
    from pyquante_amd.basis import BasisSet
    from pyquante_amd.force import UHFWavefunction, hf_force, uhf_gradient
    from pyquante_amd.hartree_fock import uhf
    mol = _oh_radical()
    E, g = uhf_gradient(mol, "6-31g**", ConvCriteria=1e-12)
    g = g.cpu().numpy()
    assert np.max(np.abs(g.sum(0))) < 1e-8
    X = mol.coords()
    h = 1e-4
    for a in range(2):
        for x in range(3):
            Xp = X.copy(); Xp[a, x] += h
            Xm = X.copy(); Xm[a, x] -= h
            ep = uhf_gradient(mol.copy_with_coords(Xp), "6-31g**", ConvCriteria=1e-12)[0]
            em = uhf_gradient(mol.copy_with_coords(Xm), "6-31g**", ConvCriteria=1e-12)[0]
            fd = (ep - em) / (2 * h)
            assert abs(fd - g[a, x]) < 1e-6, (a, x, fd, g[a, x])
    bs = BasisSet(mol, "6-31g**")
    e2, orbe, orbs = uhf(mol, bfs=bs.bfs, ConvCriteria=1e-12, MaxIter=400, reference_boys=False)
    na, nb = mol.get_alphabeta()
    f = hf_force(mol, UHFWavefunction(orbs, orbe, na, nb), "6-31g**")
    assert np.max(np.abs(f + g)) < 1e-9


def _ensemble_scf(mol, occ, basis="6-31g**", conv=1e-12, maxit=300):
    """Spin-symmetric fixed-occupation SCF (test helper): both spins put the
    fractions occ on their lowest orbitals, D = sum_i occ_i C_i C_i^T per
    spin, F = h + 2 J(D) - K(D), E = 2 tr(D h) + 2 tr(D J) - tr(D K) + V_nn.
    Returns (E, orbe, C) at convergence."""
    from pyquante_amd.basis import BasisSet
    from pyquante_amd.engine import Engine
    from pyquante_amd.hartree_fock import DIIS, chol_orth, geigh
    eng = Engine(BasisSet(mol, basis), mol)
    eng.build(1e-12, 0.0)
    S, T, V = eng.get1ints()
    h = T + V
    X = chol_orth(S)
    orbe, C = geigh(h, S, X)
    f = torch.tensor(occ, dtype=torch.float64, device=eng.device)
    n = len(occ)
    avg = DIIS(S)
    eold = 0.0
    for _ in range(maxit):
        D = (C[:, :n] * f) @ C[:, :n].T
        J, K = eng.fock_jk(D)
        E = float((D * (2 * h + 2 * J - K)).sum()) + mol.get_enuke()
        orbe, C = geigh(avg.getF(h + 2 * J - K, D), S, X)
        if abs(E - eold) < conv:
            break
        eold = E
    return E, orbe, C


def test_fixedocc_forces_integer_and_fractional_occupations():
    """force.py:91-108 with Wavefunction.py:161-199: (1) integer
    occupations of an RHF wavefunction give its RHF forces; (2) on an open
    shell, the beta energy-weighted density takes the alpha orbital energies
    (Wavefunction.py:197), so the forces equal the UHF ones of the same
    orbitals with orbe_b replaced by orbe_a; (3) fractional occupations: a
    spin-symmetric ensemble with 0.7 / 0.3 in the HOMO / LUMO of water,
    converged at each geometry, against central finite differences
    (h = 1e-4 bohr) of its energy, every coordinate; (4) hf_force dispatch."""
    from pyquante_amd.basis import BasisSet
    from pyquante_amd.engine import Engine
    from pyquante_amd.force import (FixedOccWavefunction, UHFWavefunction, Wavefunction, fixedocc_uhf_force,
                                    hf_force, rhf_force, scf, uhf_force)
    from pyquante_amd.hartree_fock import uhf
    mol = _h2o_bent()
    bs = BasisSet(mol, "6-31g**")
    eng = Engine(bs, mol)
    eng.build(1e-12, 0.0)
    E, orbe, C, D, W = scf(eng, mol, ConvCriteria=1e-12)
    nocc = mol.get_closedopen()[0]
    f_r = rhf_force(mol, Wavefunction(C, orbe, nocc), bs)
    wf = FixedOccWavefunction(C, C, orbe, orbe, [1.0] * nocc, [1.0] * nocc)
    f_x = fixedocc_uhf_force(mol, wf, bs)
    assert np.max(np.abs(f_x - f_r)) < 1e-10, np.max(np.abs(f_x - f_r))
    # open shell: the reference's beta weighting
    oh = _oh_radical()
    bo = BasisSet(oh, "6-31g**")
    _, (ea, eb), (ca, cb) = uhf(oh, bfs=bo.bfs, ConvCriteria=1e-12, MaxIter=400, reference_boys=False)
    na, nb = oh.get_alphabeta()
    assert np.max(np.abs(np.asarray(ea)[:nb] - np.asarray(eb)[:nb])) > 1e-3
    f_o = fixedocc_uhf_force(oh, FixedOccWavefunction(ca, cb, ea, eb, [1.0] * na, [1.0] * nb), bo)
    f_u = uhf_force(oh, UHFWavefunction((ca, cb), (ea, ea), na, nb), bo)
    assert np.max(np.abs(f_o - f_u)) < 1e-10
    assert np.max(np.abs(f_o - uhf_force(oh, UHFWavefunction((ca, cb), (ea, eb), na, nb), bo))) > 1e-6
    # fractional occupations against finite differences of the ensemble energy
    occ = [1.0] * (nocc - 1) + [0.7, 0.3]
    E0, orbe0, C0 = _ensemble_scf(mol, occ)
    wf = FixedOccWavefunction(C0, C0, orbe0, orbe0, occ, occ)
    g = -fixedocc_uhf_force(mol, wf, bs)
    assert np.max(np.abs(g.sum(0))) < 1e-8
    X = mol.coords()
    h = 1e-4
    for a in range(3):
        for x in range(3):
            Xp = X.copy(); Xp[a, x] += h
            Xm = X.copy(); Xm[a, x] -= h
            fd = (_ensemble_scf(mol.copy_with_coords(Xp), occ)[0]
                  - _ensemble_scf(mol.copy_with_coords(Xm), occ)[0]) / (2 * h)
            assert abs(fd - g[a, x]) < 1e-6, (a, x, fd, g[a, x])
    f = hf_force(mol, wf, "6-31g**")
    assert np.max(np.abs(f + g)) < 1e-9
