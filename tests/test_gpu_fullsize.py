"""GPU: the benchmark configurations at full size, checked through properties
that do not need the oracle at that size.

- (H2O)_8 6-31G** unscreened: 4,656 shell pairs, so 4656*4657/2 = 10,841,496
  unique shell quartets and 20100*20101/2 = 202,015,050 unique function
  quartets (closed forms of the enumeration in Ints.py:146-167);
- J and K are linear in D: J(D1 + 2 D2) = J(D1) + 2 J(D2) to 1e-10 relative,
  with screening and primitive screening on (they depend on the pairs, not on
  D);
- the two-density pass equals two single builds to 1e-11;
- (H2O)_8: every one of the 202,015,050 unique ERIs by the shell-quartet HGP
  kernels (primitive-split tiles adding partial blocks with atomics, all
  classes in one launch) equals the Rys store array to 1e-12, and the J/K
  digested from that array equals the direct Fock build to 1e-10;
- (H2O)_16 6-31G** (400 functions): the mixed-precision energy functional
  tr(D(2J - K)) is within 1e-6 Eh of the FP64 one for a seeded density (the
  accuracy bound of the mixed path).
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _engine(lattice, seed=0):
    from pyquante_amd.basis import BasisSet
    from pyquante_amd.engine import Engine
    from pyquante_amd.molecule import water_cluster
    mol = water_cluster(lattice, seed=seed)
    return Engine(BasisSet(mol, "6-31g**"), mol)


def _density(n, k, seed):
    rng = np.random.default_rng(seed)
    C = rng.standard_normal((n, k)) * 0.1
    return torch.as_tensor(C @ C.T, device="cuda")


def test_h2o8_counts_and_linearity():
    eng = _engine((2, 2, 2))
    assert eng.nbf == 200
    eng.build(0.0)
    st = eng.stats()
    assert st["quartets_fp64"] == 4656 * 4657 // 2
    assert st["fquartets_fp64"] == 20100 * 20101 // 2
    eng.build(1e-12)
    D1 = _density(eng.nbf, 20, 1)
    D2 = _density(eng.nbf, 20, 2)
    J1, K1 = eng.fock_jk(D1)
    J2, K2 = eng.fock_jk(D2)
    J3, K3 = eng.fock_jk(D1 + 2.0 * D2)
    for a, b in ((J3, J1 + 2.0 * J2), (K3, K1 + 2.0 * K2)):
        assert torch.max(torch.abs(a - b)).item() < 1e-10 * torch.max(torch.abs(b)).item()
    (Ja, Ka), (Jb, Kb) = eng.fock_jk_multi([D1, D2])
    for a, b in ((Ja, J1), (Ka, K1), (Jb, J2), (Kb, K2)):
        assert torch.max(torch.abs(a - b)).item() < 1e-11


def test_h2o16_mixed_precision_energy_bound():
    from pyquante_amd import _abi
    eng = _engine((2, 2, 4))
    assert eng.nbf == 400
    D = _density(eng.nbf, 40, 3)
    eng.build(1e-12, 0.0)
    J64, K64 = eng.fock_jk(D)
    eng.build(1e-12, 1e-4)
    assert eng.stats()["fquartets_fp32"] > eng.stats()["fquartets_fp64"]
    J32, K32 = eng.fock_jk(D, precision=_abi.PQ_PREC_MIXED)
    e64 = float(torch.sum(D * (2.0 * J64 - K64)))
    e32 = float(torch.sum(D * (2.0 * J32 - K32)))
    print("(H2O)_16 mixed - FP64 energy functional: %.2e" % (e32 - e64), flush=True)
    assert abs(e32 - e64) < 1e-6


def test_h2o8_hgp_full_array_matches_rys():
    eng = _engine((2, 2, 2))
    hgp = eng.get2ints("hgp", 0)
    assert hgp.numel() == 202015050
    assert bool(torch.isfinite(hgp).all())
    rys = eng.get2ints("rys")
    d = float((hgp - rys).abs().max())
    del rys
    print("(H2O)_8 HGP vs Rys, all unique ERIs: max |d| = %.2e" % d, flush=True)
    assert d < 1e-12
    # J/K digested from the HGP array against the direct build
    eng.build(0.0, 0.0)
    D = _density(eng.nbf, 20, 5)
    from pyquante_amd.engine import digest_stored
    Js, Ks = digest_stored(hgp, D)
    del hgp
    J, K = eng.fock_jk(D)
    scale = float(J.abs().max())
    assert float((Js - J).abs().max()) < 1e-10 * scale
    assert float((Ks - K).abs().max()) < 1e-10 * scale


def test_lds_limit_past_680_functions():
    """(H2O)_28 6-31G** (700 functions): the d-bra classes need more than
    64 KB for their LDS K rows.  With the device's limit (queried, 160 KB on
    gfx950) they keep them in LDS; lowered to 64 KB they fall back to global
    atomics.  The two accumulation paths give the same J and K."""
    eng = _engine((2, 2, 7))
    assert eng.nbf == 700
    dev = eng.set_lds_limit(0)
    print("device LDS per workgroup: %d bytes" % dev, flush=True)
    assert dev >= 65536
    eng.build(1e-12)
    D = _density(eng.nbf, 60, 11)
    J1, K1 = eng.fock_jk(D)
    eng.set_lds_limit(65536)
    J2, K2 = eng.fock_jk(D)
    eng.set_lds_limit(0)
    scale = float(J1.abs().max())
    dj = float((J1 - J2).abs().max())
    dk = float((K1 - K2).abs().max())
    print("(H2O)_28: LDS K rows vs global atomics: max |dJ| %.2e, |dK| %.2e (scale %.1f)" % (dj, dk, scale), flush=True)
    assert dj < 1e-11 * scale and dk < 1e-11 * scale
