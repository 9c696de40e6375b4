"""GPU: the deterministic J/K build (pq_engine_set_deterministic).

Every contribution is accumulated as exact integer limbs, so for one launch
configuration J and K must be bit-identical (torch.equal) whatever the order
of execution: repeated builds, 0-4 side streams, graph replays and eager work
between them.  Bra-run caps and the LDS limit choose between kernels
(k_eri_group, k_eri_lane with LDS or global accumulators) that the compiler
may contract into FMAs differently, so across those the integrals, and J/K,
may differ in the last bits: each such configuration must be bit-identical
with itself and within 1e-12 of the others.  Everything agrees with the
FP64-atomic build and the stored integrals digested on the GPU to 1e-11
absolute.  The oracle pins the plain build in test_gpu_eri.py; the limb
arithmetic is pinned on the CPU in test_fixed_point.py."""
import numpy as np
import pytest
import torch

from pyquante_amd._abi import PQError

pytestmark = pytest.mark.gpu


def _engine(lattice, seed):
    from pyquante_amd.basis import BasisSet
    from pyquante_amd.engine import Engine
    from pyquante_amd.molecule import water_cluster
    mol = water_cluster(lattice, seed=seed)
    return Engine(BasisSet(mol, "6-31g**"), mol)


def _density(n, seed):
    rng = np.random.default_rng(seed)
    C = rng.standard_normal((n, 20)) * 0.1
    return torch.as_tensor(C @ C.T, device="cuda")


def _close(a, b, tol=1e-11):
    return torch.max(torch.abs(a - b)).item() < tol


def test_deterministic_matches_stored_and_atomic_builds():
    from pyquante_amd.engine import digest_stored
    eng = _engine((2, 1, 1), 1)
    D = _density(eng.nbf, 11)
    Js, Ks = digest_stored(eng.get2ints(), D)
    eng.build(0.0, 0.0)
    Ja, Ka = eng.fock_jk(D)
    eng.set_deterministic(True)
    Jd, Kd = eng.fock_jk(D)
    print("deterministic vs stored: %.2e %.2e" % (torch.max(torch.abs(Jd - Js)).item(),
                                                 torch.max(torch.abs(Kd - Ks)).item()), flush=True)
    assert _close(Jd, Js) and _close(Kd, Ks)
    assert _close(Jd, Ja) and _close(Kd, Ka)
    # symmetric by construction (J = Jacc + Jacc^T from the same integers)
    assert torch.equal(Jd, Jd.T) and torch.equal(Kd, Kd.T)
    eng.set_deterministic(False)
    J2, K2 = eng.fock_jk(D)
    assert _close(J2, Ja, 1e-12) and _close(K2, Ka, 1e-12)


def test_bit_identical_whatever_the_execution_order():
    from pyquante_amd import _abi
    eng = _engine((2, 2, 1), 2)
    eng.build(1e-12, 0.0)
    D = _density(eng.nbf, 12)
    eng.set_deterministic(True)
    S = _abi.PQ_CAP_STAGE_D
    configs = [("default caps", lambda: None), ("cap 0", lambda: eng.set_group_cap(0)),
               ("cap 4", lambda: eng.set_group_cap(4)), ("cap 16 + D rows", lambda: eng.set_group_cap(16 | S)),
               ("K rows in global memory", lambda: eng.set_lds_limit(4096))]
    orders = [("again", lambda: None), ("1 stream", lambda: eng.set_streams(1)),
              ("no side streams", lambda: eng.set_streams(0)), ("4 streams", lambda: eng.set_streams(4))]
    first = None
    for cname, config in configs:
        config()
        J0, K0 = eng.fock_jk(D)
        for oname, order in orders:
            order()
            J, K = eng.fock_jk(D)
            print("deterministic: %s, %s" % (cname, oname), flush=True)
            assert torch.equal(J, J0) and torch.equal(K, K0), (cname, oname)
        if first is None:
            first = (J0, K0)
        else:
            assert _close(J0, first[0], 1e-12) and _close(K0, first[1], 1e-12), cname
    eng.set_lds_limit(0)


def test_graph_replays_bit_identical():
    eng = _engine((2, 1, 1), 3)
    eng.build(1e-12, 0.0)
    eng.set_group_cap(4)
    eng.set_deterministic(True)
    D1 = _density(eng.nbf, 13)
    D2 = _density(eng.nbf, 14)
    J1, K1 = eng.fock_jk(D1)
    J2, K2 = eng.fock_jk(D2)
    g = eng.capture_fock_jk(D1)
    for rep in range(4):
        print("deterministic graph replay %d" % rep, flush=True)
        D, Jr, Kr = (D1, J1, K1) if rep % 2 == 0 else (D2, J2, K2)
        J, K = g(D)
        torch.cuda.synchronize()
        assert torch.equal(J, Jr) and torch.equal(K, Kr)
        Je, Ke = eng.fock_jk(D2)   # eager work between replays
        torch.cuda.synchronize()
        assert torch.equal(Je, J2) and torch.equal(Ke, K2)


def test_h2o8_deterministic_against_atomic_build():
    """The benchmark molecule: bit-identical repeats, within 1e-11 of the
    FP64-atomic build, and the RHF energy functional to 1e-10 Eh."""
    eng = _engine((2, 2, 2), 0)
    eng.build(1e-12, 0.0)
    D = _density(eng.nbf, 15)
    Ja, Ka = eng.fock_jk(D)
    eng.set_deterministic(True)
    Jd, Kd = eng.fock_jk(D)
    for _ in range(2):
        J, K = eng.fock_jk(D)
        assert torch.equal(J, Jd) and torch.equal(K, Kd)
    assert _close(Jd, Ja) and _close(Kd, Ka)
    ea = torch.sum(D * (2 * Ja - Ka)).item()
    ed = torch.sum(D * (2 * Jd - Kd)).item()
    assert abs(ea - ed) < 1e-10


def test_batch_engine_deterministic():
    from pyquante_amd.batch import BatchEngine, seeded_water_batch
    be = BatchEngine(seeded_water_batch(nmol=6, nmax=3, seed=9))
    be.build(1e-12, 0.0)
    rng = np.random.default_rng(4)
    blocks = []
    for n in be.mol_nbf:
        C = rng.standard_normal((n, 6)) * 0.1
        blocks.append((C @ C.T).reshape(-1))
    D = torch.as_tensor(np.concatenate(blocks), device="cuda")
    Ja, Ka = be.fock_jk(D)
    be.set_deterministic(True)
    J0, K0 = be.fock_jk(D)
    assert _close(J0, Ja) and _close(K0, Ka)
    be.set_streams(1)
    J1, K1 = be.fock_jk(D)
    assert torch.equal(J1, J0) and torch.equal(K1, K0)


def test_deterministic_two_densities_and_mixed_precision():
    """The deterministic build of two densities in one pass and of
    mixed-precision work lists (FP32 quartets, the same limb accumulation):
    bit-identical from call to call and across 0, 1 and 4 side streams; the
    two-density pass within 1e-12 of two single deterministic builds and
    1e-11 of the FP64-atomic pass; the mixed build within 1e-10 of the
    atomic mixed build, with its energy functional within 1e-6 Eh of the
    FP64 one."""
    from pyquante_amd import _abi
    eng = _engine((2, 1, 1), 5)
    eng.build(1e-12, 0.0)
    Da, Db = _density(eng.nbf, 16), _density(eng.nbf, 17)
    atomic = eng.fock_jk_multi([Da, Db])
    eng.set_deterministic(True)
    single = [eng.fock_jk(Da), eng.fock_jk(Db)]
    ref = eng.fock_jk_multi([Da, Db])
    for n in (0, 1, 4):
        eng.set_streams(n)
        for _ in range(2):
            got = eng.fock_jk_multi([Da, Db])
            for (J, K), (J0, K0) in zip(got, ref):
                assert torch.equal(J, J0) and torch.equal(K, K0)
    eng.set_streams(4)
    for (J, K), (Js, Ks), (Ja, Ka) in zip(ref, single, atomic):
        assert torch.equal(J, J.T) and torch.equal(K, K.T)
        assert _close(J, Js, 1e-12) and _close(K, Ks, 1e-12)
        assert _close(J, Ja) and _close(K, Ka)
    # mixed precision
    E64 = float((Da * (2 * single[0][0] - single[0][1])).sum())
    eng.set_deterministic(False)
    eng.build(1e-12, 1e-4)
    assert eng.stats()["quartets_fp32"] > 0
    Jm, Km = eng.fock_jk(Da, precision=_abi.PQ_PREC_MIXED)
    eng.set_deterministic(True)
    Jd, Kd = eng.fock_jk(Da, precision=_abi.PQ_PREC_MIXED)
    for n in (0, 1, 4):
        eng.set_streams(n)
        J, K = eng.fock_jk(Da, precision=_abi.PQ_PREC_MIXED)
        assert torch.equal(J, Jd) and torch.equal(K, Kd)
    print("deterministic mixed vs atomic mixed: %.2e %.2e" % (torch.max(torch.abs(Jd - Jm)).item(),
                                                            torch.max(torch.abs(Kd - Km)).item()), flush=True)
    assert _close(Jd, Jm, 1e-10) and _close(Kd, Km, 1e-10)
    Em = float((Da * (2 * Jd - Kd)).sum())
    assert abs(Em - E64) < 1e-6, (Em, E64)
