"""GPU: launch-shape variants of the direct Fock build give the same J/K.

Bra runs (one block walking several tiles of one bra shell pair), the
autotuner that picks their length per class, and HIP-graph replay of a whole
build must all reproduce the plain launch to within the non-associativity of
the FP64 atomic sums (1e-11 absolute on J/K of order 1).  The plain launch
itself is pinned against the oracle in test_gpu_eri.py."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _engine(lattice, seed):
    from pyquante_amd.basis import BasisSet
    from pyquante_amd.engine import Engine
    from pyquante_amd.molecule import water_cluster
    mol = water_cluster(lattice, seed=seed)
    eng = Engine(BasisSet(mol, "6-31g**"), mol)
    return eng


def _density(n, seed):
    rng = np.random.default_rng(seed)
    C = rng.standard_normal((n, 20)) * 0.1
    return torch.as_tensor(C @ C.T, device="cuda")


def test_bra_runs_match_plain_launch():
    from pyquante_amd import _abi
    from pyquante_amd.engine import digest_stored
    eng = _engine((2, 1, 1), 1)
    D = _density(eng.nbf, 11)
    Js, Ks = digest_stored(eng.get2ints(), D)
    for cut, fp32 in ((0.0, 0.0), (1e-12, 1e-4)):
        prec = _abi.PQ_PREC_MIXED if fp32 > 0 else _abi.PQ_PREC_FP64
        eng.build(cut, fp32)
        eng.set_group_cap(0)
        J0, K0 = eng.fock_jk(D, precision=prec)
        if fp32 == 0.0:
            assert torch.max(torch.abs(J0 - Js)).item() < 1e-11
            assert torch.max(torch.abs(K0 - Ks)).item() < 1e-11
        S = _abi.PQ_CAP_STAGE_D
        for cap in (1, 3, 4, 16, 1024, 1 | S, 4 | S, 1024 | S):
            eng.set_group_cap(cap)
            J, K = eng.fock_jk(D, precision=prec)
            print("bra runs: cut %g fp32 %g cap %d" % (cut, fp32, cap), flush=True)
            assert torch.max(torch.abs(J - J0)).item() < 1e-11
            assert torch.max(torch.abs(K - K0)).item() < 1e-11
    # a rebuild keeps the cap that was set
    eng.set_group_cap(4)
    eng.build(0.0, 0.0)
    J, K = eng.fock_jk(D)
    assert torch.max(torch.abs(K - Ks)).item() < 1e-11


def test_autotune_choices_and_results():
    from pyquante_amd import _abi
    eng = _engine((2, 2, 1), 2)
    D = _density(eng.nbf, 12)
    eng.build(1e-12, 1e-4)
    eng.set_group_cap(0)
    J0, K0 = eng.fock_jk(D, precision=_abi.PQ_PREC_MIXED)
    caps = eng.autotune(D, reps=1)
    assert len(caps) == 21
    S = _abi.PQ_CAP_STAGE_D
    assert all(c in (0, 4, 8, 16, 32, 4 | S, 8 | S, 16 | S, 32 | S) for pair in caps for c in pair)
    J, K = eng.fock_jk(D, precision=_abi.PQ_PREC_MIXED)
    assert torch.max(torch.abs(J - J0)).item() < 1e-11
    assert torch.max(torch.abs(K - K0)).item() < 1e-11


def test_whole_build_tuning():
    """The whole-build refinement returns a valid table that the engine keeps,
    and builds with it give the plain launch's J/K."""
    from pyquante_amd import _abi
    eng = _engine((2, 1, 1), 5)
    D = _density(eng.nbf, 15)
    eng.build(1e-12, 0.0)
    eng.set_group_cap(0)
    J0, K0 = eng.fock_jk(D)
    caps, ms = eng.autotune_build(D, reps=1, steps=2)
    assert ms > 0 and len(caps) == 21
    S = _abi.PQ_CAP_STAGE_D
    assert all(c in (-1, 0, 4, 8, 16, 32, 4 | S, 8 | S, 16 | S, 32 | S) for pair in caps for c in pair)
    J, K = eng.fock_jk(D)
    assert torch.max(torch.abs(J - J0)).item() < 1e-11
    assert torch.max(torch.abs(K - K0)).item() < 1e-11


@pytest.mark.parametrize("streams", [0, 4])
def test_graph_replay_matches_eager(streams):
    eng = _engine((2, 1, 1), 3)
    eng.set_streams(streams)
    eng.build(1e-12, 0.0)
    eng.set_group_cap(4)
    D1 = _density(eng.nbf, 13)
    D2 = _density(eng.nbf, 14)
    J1, K1 = eng.fock_jk(D1)
    J2, K2 = eng.fock_jk(D2)
    g = eng.capture_fock_jk(D1)
    for rep in range(4):
        print("graph replay %d streams %d" % (rep, streams), flush=True)
        D, Jr, Kr = (D1, J1, K1) if rep % 2 == 0 else (D2, J2, K2)
        J, K = g(D)
        torch.cuda.synchronize()
        assert torch.max(torch.abs(J - Jr)).item() < 1e-11
        assert torch.max(torch.abs(K - Kr)).item() < 1e-11
        # eager work between replays
        Je, _ = eng.fock_jk(D2)
        torch.cuda.synchronize()
        assert torch.max(torch.abs(Je - J2)).item() < 1e-11


def test_scf_with_graph_replayed_fock_builds():
    from pyquante_amd.force import scf
    eng = _engine((2, 1, 1), 5)
    eng.build(1e-12, 0.0)
    from pyquante_amd.molecule import water_cluster
    mol = water_cluster((2, 1, 1), seed=5)
    e0 = scf(eng, mol, ConvCriteria=1e-10)[0]
    e1 = scf(eng, mol, ConvCriteria=1e-10, graph=True)[0]
    assert abs(e0 - e1) < 1e-9



def test_fused_light_class_launch_matches_per_class():
    """One launch over the light classes' bra runs (pq_engine_set_fused)
    gives the per-class launches' J and K, eager and graph-replayed."""
    import numpy as np
    import torch
    from pyquante_amd import tuning
    from pyquante_amd.basis import BasisSet
    from pyquante_amd.engine import Engine
    from pyquante_amd.molecule import water_cluster
    mol = water_cluster((2, 2, 2), seed=0)
    eng = Engine(BasisSet(mol, "6-31g**"), mol)
    eng.build(1e-12)
    caps = tuning.load("h2o8_631ss_fp64")
    if caps is not None:
        eng.set_group_caps(caps)
    else:
        eng.set_group_cap(16)
    rng = np.random.default_rng(23)
    C = rng.standard_normal((eng.nbf, 20)) * 0.1
    D = torch.as_tensor(C @ C.T, device=eng.device)
    J0, K0 = eng.fock_jk(D)
    eng.set_fused(True)
    J1, K1 = eng.fock_jk(D)
    g = eng.capture_fock_jk(D)
    for _ in range(3):
        J2, K2 = g(D)
        torch.cuda.synchronize()
        for a, b in ((J1, J0), (K1, K0), (J2, J0), (K2, K0)):
            assert torch.max(torch.abs(a - b)).item() < 1e-11
    eng.set_fused(False)
