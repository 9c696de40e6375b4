"""GPU: two densities digested in one pass (the unrestricted J/K path)
against two single-density builds, and the UHF driver against the
reference's golden energies."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _density(n, seed, k=12):
    rng = np.random.default_rng(seed)
    C = rng.standard_normal((n, k)) * 0.1
    return torch.as_tensor(C @ C.T, device="cuda")


@pytest.mark.parametrize("cap", [0, 4, 4 | 4096])
def test_two_density_pass_matches_single_builds(cap):
    from pyquante_amd.basis import BasisSet
    from pyquante_amd.engine import Engine
    from pyquante_amd.molecule import water_cluster
    mol = water_cluster((2, 1, 1), seed=6)
    eng = Engine(BasisSet(mol, "6-31g**"), mol)
    eng.build(1e-12, 0.0)
    eng.set_group_cap(cap)
    Da, Db = _density(eng.nbf, 21), _density(eng.nbf, 22, 9)
    Ja, Ka = eng.fock_jk(Da)
    Jb, Kb = eng.fock_jk(Db)
    (Ja2, Ka2), (Jb2, Kb2) = eng.fock_jk_multi([Da, Db])
    for x, y in ((Ja, Ja2), (Ka, Ka2), (Jb, Jb2), (Kb, Kb2)):
        assert torch.max(torch.abs(x - y)).item() < 1e-11
    (J1, K1), = eng.fock_jk_multi([Da])
    assert torch.max(torch.abs(J1 - Ja)).item() < 1e-11


def test_uhf_no_reference_energy():
    """Tests/no_uhf.py:13-23: NO at 2.12955 bohr, doublet, 6-31G**, energy
    -129.247769 within 1e-4 (the reference test's tolerance)."""
    from pyquante_amd.hartree_fock import uhf
    from pyquante_amd.molecule import Molecule
    mol = Molecule("NO", [(7, (0, 0, 0)), (8, (2.12955, 0, 0))], multiplicity=2)
    e, _, _ = uhf(mol)
    print("NO UHF %.8f" % e, flush=True)
    assert abs(e - (-129.247769)) < 1e-4


def test_uhf_closed_shell_equals_rhf():
    from pyquante_amd.hartree_fock import rhf, uhf
    from pyquante_amd.molecule import Molecule
    r = 1 / 0.52918
    mol = Molecule("h2o", [(8, (0, 0, 0)), (1, (r, 0, 0)), (1, (0, r, 0))])
    e_r, _, _ = rhf(mol, basis_data="6-31g**", ConvCriteria=1e-11, MaxIter=100, DoAveraging=True)
    e_u, _, _ = uhf(mol, basis_data="6-31g**", ConvCriteria=1e-11, MaxIter=200)
    assert abs(e_u - e_r) < 1e-8
