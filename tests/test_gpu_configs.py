"""GPU: the benchmark configurations at their own sizes.

C2  (H2O)_8 6-31G** (200 functions): 20,000 seeded unique function quartets
    of pq_get2ints against the compiled reference crys.c (fixture), and the
    RHF energy through the direct Rys build against RHF through the stored
    HGP array (two independent integral algorithms);
C3  (H2O)_16 (400 functions): the mixed-precision RHF energy against the
    FP64 one;
C4  (H2O)_8 forces: the analytic gradient against central finite
    differences (h = 1e-4 bohr) on 6 seeded coordinates, translational
    invariance (sum of forces) and rotational invariance (net torque);
C5  the 512-molecule batch (seeded (H2O)_n, n uniform in 1..12): mixed
    against FP64 per molecule, and molecules spanning n = 1..12 against
    single-molecule engines.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _h2o8():
    from pyquante_amd.molecule import water_cluster
    return water_cluster((2, 2, 2), seed=0)


def test_c2_h2o8_integrals_vs_compiled_crys(golden):
    from pyquante_amd.basis import BasisSet
    from pyquante_amd.engine import Engine
    from pyquante_amd.Ints import ijkl2intindex
    g = golden("h2o8_631ss")
    mol = _h2o8()
    # the seeded generator still gives the fixture's nuclei
    assert np.max(np.abs(mol.coords() - g["xyz"])) < 1e-12
    eng = Engine(BasisSet(mol, "6-31g**"), mol)
    ints = eng.get2ints()
    idx = torch.as_tensor([ijkl2intindex(*map(int, q)) for q in g["ijkl"]], device=ints.device)
    v = ints[idx].cpu().numpy()
    del ints
    d = np.abs(v - g["crys"])
    print("C2: 20,000 (H2O)_8 quartets vs compiled crys: max |d| %.2e (max |v| %.2f)"
          % (d.max(), np.abs(g["crys"]).max()), flush=True)
    assert d.max() < 1e-12


def test_c2_h2o8_rhf_rys_direct_vs_hgp_stored():
    from pyquante_amd import Ints
    from pyquante_amd.hartree_fock import rhf
    mol = _h2o8()
    e_r = rhf(mol, basis_data="6-31g**", ConvCriteria=1e-10, MaxIter=60, DoAveraging=True)[0]
    old = Ints.hgp_compat_boys
    Ints.hgp_compat_boys = False
    try:
        e_h = rhf(mol, basis_data="6-31g**", ConvCriteria=1e-10, MaxIter=60, DoAveraging=True, method="hgp")[0]
    finally:
        Ints.hgp_compat_boys = old
    print("C2: (H2O)_8 RHF Rys direct %.10f, HGP stored %.10f" % (e_r, e_h), flush=True)
    assert abs(e_r - e_h) < 1e-9
    assert -610.0 < e_r < -600.0  # 8 waters at about -76.0 Eh each


def test_c3_h2o16_mixed_rhf_energy():
    from pyquante_amd import _abi
    from pyquante_amd.basis import BasisSet
    from pyquante_amd.engine import Engine
    from pyquante_amd.force import scf
    from pyquante_amd.molecule import water_cluster
    mol = water_cluster((2, 2, 4), seed=0)
    eng = Engine(BasisSet(mol, "6-31g**"), mol)
    eng.build(1e-12, 0.0)
    e64 = scf(eng, mol, ConvCriteria=1e-10)[0]
    eng.build(1e-12, 1e-4)
    st = eng.stats()
    frac = st["fquartets_fp32"] / (st["fquartets_fp32"] + st["fquartets_fp64"])
    emx = scf(eng, mol, ConvCriteria=1e-10, precision=_abi.PQ_PREC_MIXED)[0]
    print("C3: (H2O)_16 RHF FP64 %.10f, mixed %.10f, |dE| %.2e Eh (%.1f%% of function quartets in FP32)"
          % (e64, emx, abs(emx - e64), 100 * frac), flush=True)
    assert frac > 0.5
    assert abs(emx - e64) < 1e-6


def test_c4_h2o8_forces_finite_differences_and_invariances():
    from pyquante_amd.force import rhf_gradient
    mol = _h2o8()
    E, g = rhf_gradient(mol, "6-31g**", ConvCriteria=1e-11)
    g = g.cpu().numpy()
    X = mol.coords()
    F = -g
    net = F.sum(0)
    torque = np.cross(X, F).sum(0)
    print("C4: E %.10f, |sum F| %.1e, |torque| %.1e" % (E, np.abs(net).max(), np.abs(torque).max()), flush=True)
    assert np.abs(net).max() < 1e-7
    assert np.abs(torque).max() < 1e-6
    rng = np.random.default_rng(12)
    h = 1e-4
    picks = set()
    while len(picks) < 6:
        picks.add((int(rng.integers(len(mol))), int(rng.integers(3))))
    for a, x in sorted(picks):
        Xp = X.copy(); Xp[a, x] += h
        Xm = X.copy(); Xm[a, x] -= h
        ep = rhf_gradient(mol.copy_with_coords(Xp), "6-31g**", ConvCriteria=1e-11)[0]
        em = rhf_gradient(mol.copy_with_coords(Xm), "6-31g**", ConvCriteria=1e-11)[0]
        fd = (ep - em) / (2 * h)
        print("  atom %2d %s: analytic %+.9f, central FD %+.9f, |d| %.1e" % (a, "xyz"[x], g[a, x], fd,
                                                                          abs(fd - g[a, x])), flush=True)
        assert abs(fd - g[a, x]) < 1e-6


def test_c5_full_batch_mixed_vs_fp64_and_single_engines():
    from pyquante_amd import _abi
    from pyquante_amd.basis import BasisSet
    from pyquante_amd.batch import BatchEngine, batch_forces, batch_rhf, seeded_water_batch
    from pyquante_amd.engine import Engine
    from pyquante_amd.force import scf
    mols = seeded_water_batch(nmol=512, nmax=12, seed=0)
    ns = [len(list(m)) // 3 for m in mols]
    assert len(mols) == 512 and min(ns) == 1 and max(ns) == 12
    be = BatchEngine(mols)
    assert be.nmax == 300
    be.build(1e-12, 0.0)
    E64, _, _, D64, W64, it64 = batch_rhf(be, ConvCriteria=1e-10)
    F64 = batch_forces(be, D64, W64)
    be.build(1e-12, 1e-4)
    E32, _, _, D32, W32, it32 = batch_rhf(be, ConvCriteria=1e-10, precision=_abi.PQ_PREC_MIXED)
    F32 = batch_forces(be, D32, W32)
    dE = float((E32 - E64).abs().max())
    dF = max(float((a - b).abs().max()) for a, b in zip(F32, F64))
    print("C5: 512 molecules, %d functions: SCF %d / %d iterations; mixed vs FP64 max |dE| %.1e Eh, "
          "max |dF| %.1e Eh/bohr" % (be.nbf, it64, it32, dE, dF), flush=True)
    assert dE < 1e-6 and dF < 1e-6
    # molecules of every size class against their own single-molecule engines
    for n in (1, 2, 3, 5, 7, 9, 11, 12):
        m = ns.index(n)
        eng = Engine(BasisSet(mols[m], "6-31g**"), mols[m])
        eng.build(1e-12, 0.0)
        e1, _, _, D1, W1 = scf(eng, mols[m], ConvCriteria=1e-11)
        g1 = eng.gradient(D1, W1)
        print("  molecule %3d (n = %2d): E batch %.9f single %.9f, max |dF| %.1e"
              % (m, n, float(E64[m]), e1, float((F64[m] - g1).abs().max())), flush=True)
        assert abs(float(E64[m]) - e1) < 1e-8
        assert float((F64[m] - g1).abs().max()) < 1e-6
