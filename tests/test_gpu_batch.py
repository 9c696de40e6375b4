"""GPU: a batch of independent molecules through one engine (packed
per-molecule matrices, one launch per class over all molecules) against each
molecule on its own engine."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _mols():
    from pyquante_amd.batch import seeded_water_batch
    return seeded_water_batch(nmol=5, nmax=3, seed=7)


def test_batch_fock_energies_forces_match_single_engines():
    from pyquante_amd.basis import BasisSet
    from pyquante_amd.batch import BatchEngine, batch_forces, batch_rhf
    from pyquante_amd.engine import Engine
    from pyquante_amd.force import scf
    mols = _mols()
    be = BatchEngine(mols)
    be.build(1e-12, 0.0)
    # one-electron blocks and a Fock build on a random packed density
    S, T, V = be.get1ints()
    rng = np.random.default_rng(3)
    Dblocks = []
    for n in be.mol_nbf:
        C = rng.standard_normal((n, 6)) * 0.1
        Dblocks.append(C @ C.T)
    D = torch.as_tensor(np.concatenate([b.reshape(-1) for b in Dblocks]), device="cuda")
    J, K = be.fock_jk(D)
    # bra runs with staged D rows over the packed molecule blocks
    from pyquante_amd import _abi
    be.set_group_cap(4 | _abi.PQ_CAP_STAGE_D)
    Jr, Kr = be.fock_jk(D)
    be.set_group_cap(-1)
    assert torch.max(torch.abs(Jr - J)).item() < 1e-11
    assert torch.max(torch.abs(Kr - K)).item() < 1e-11
    for m, mol in enumerate(mols):
        eng = Engine(BasisSet(mol, "6-31g**"), mol)
        eng.build(1e-12, 0.0)
        Sm, Tm, Vm = eng.get1ints()
        for x, y in ((S, Sm), (T, Tm), (V, Vm)):
            assert torch.max(torch.abs(be.blocks(x)[m] - y)).item() < 1e-12
        Jm, Km = eng.fock_jk(torch.as_tensor(Dblocks[m], device="cuda"))
        assert torch.max(torch.abs(be.blocks(J)[m] - Jm)).item() < 1e-11
        assert torch.max(torch.abs(be.blocks(K)[m] - Km)).item() < 1e-11
    # SCF and forces of every molecule at once
    E, orbe, C, Dp, Wp, it = batch_rhf(be, ConvCriteria=1e-10)
    F = batch_forces(be, Dp, Wp)
    for m, mol in enumerate(mols):
        eng = Engine(BasisSet(mol, "6-31g**"), mol)
        eng.build(1e-12, 0.0)
        e1, _, _, D1, W1 = scf(eng, mol, ConvCriteria=1e-10)
        g1 = eng.gradient(D1, W1)
        print("molecule %d (%d atoms): E batch %.10f single %.10f" % (m, len(list(mol)), float(E[m]), e1),
              flush=True)
        assert abs(float(E[m]) - e1) < 1e-8
        assert torch.max(torch.abs(F[m] - g1)).item() < 1e-6


def test_batch_mixed_precision_error():
    from pyquante_amd import _abi
    from pyquante_amd.batch import BatchEngine, batch_rhf
    mols = _mols()
    be = BatchEngine(mols)
    be.build(1e-12, 0.0)
    E64 = batch_rhf(be, ConvCriteria=1e-10)[0]
    be.build(1e-12, 1e-4)
    E32 = batch_rhf(be, ConvCriteria=1e-10, precision=_abi.PQ_PREC_MIXED)[0]
    assert float(torch.max(torch.abs(E64 - E32))) < 1e-6
