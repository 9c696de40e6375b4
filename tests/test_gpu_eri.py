"""GPU parity: the HIP engine against the oracle and the compiled reference's
golden vectors.  Tolerances: FP64 integrals 1e-12 absolute (the reference's
own roots are ~1e-13 relative), Boys 1e-13 relative, J/K 1e-10 absolute."""
import math

import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _h2o_631ss():
    from pyquante_amd.molecule import Molecule
    r = 1 / 0.52918
    return Molecule("h2o", [(8, (0, 0, 0)), (1, (r, 0, 0)), (1, (0, r, 0))])


def test_rys_roots_table_vs_oracle():
    from pyquante_amd import _abi
    rng = np.random.default_rng(0)
    X = np.concatenate([[0.0, 1e-9, 0.5, 1.0, 36.999, 37.0, 63.99, 64.0, 1e3, 1e6],
                        rng.uniform(0, 80, 400)])
    for nr in range(1, 6):
        u = np.zeros(len(X) * nr)
        w = np.zeros(len(X) * nr)
        _abi.check(_abi.lib.pq_rys_roots_batch(len(X), nr, _abi.pd(X), _abi.pd(u), _abi.pd(w)))
        u = u.reshape(-1, nr)
        w = w.reshape(-1, nr)
        for t, x in enumerate(X):
            uo, wo = O.rys_roots(nr, x)
            assert np.allclose(u[t], uo, rtol=5e-14, atol=0), (nr, x)
            assert np.allclose(w[t], wo, rtol=5e-14, atol=0), (nr, x)


def test_boys_batch(golden):
    from pyquante_amd import _abi
    g = golden("boys_nr")
    m, x = np.meshgrid(g["m"], g["x"], indexing="ij")
    m = np.ascontiguousarray(m.ravel())
    x = np.ascontiguousarray(x.ravel())
    out = np.zeros_like(m)
    _abi.check(_abi.lib.pq_boys_batch(len(m), _abi.pd(m), _abi.pd(x), 1, _abi.pd(out)))
    assert np.allclose(out, g["F"].ravel(), rtol=1e-12, atol=0)
    _abi.check(_abi.lib.pq_boys_batch(len(m), _abi.pd(m), _abi.pd(x), 0, _abi.pd(out)))
    ref = np.array([O.boys(a, b) for a, b in zip(m, x)])
    mask = x > 1e-8  # the reference clamps x to 1e-8
    assert np.allclose(out[mask], ref[mask], rtol=2e-14, atol=0)
    from pyquante_amd import cints
    assert abs(cints.Fgamma(0, 1) - 0.7468241328124271) < 1e-8      # TestMath.c:19
    assert abs(cints.Fgamma(0.5, 6.8) - 0.073447516533246701) < 1e-8  # TestMath.c:20


def test_prim_classes_vs_compiled_crys(golden):
    from pyquante_amd._compat import contr_coulomb_batch
    g = golden("prim_classes")
    P, Org, A = g["powers"], g["origins"], g["alphas"]
    quartets = [tuple(([A[t, q]], [1.0], [1.0], Org[t, 3 * q:3 * q + 3], P[t, 3 * q:3 * q + 3])
                      for q in range(4)) for t in range(len(P))]
    v = contr_coulomb_batch(quartets)
    assert np.allclose(v, g["crys"], rtol=1e-12, atol=1e-14)


def test_get2ints_h2o_sto3g(golden):
    from pyquante_amd.Ints import get2ints
    from pyquante_amd.basis import BasisSet
    from pyquante_amd.molecule import h2o_sto3g_config
    mol = h2o_sto3g_config()
    ints = get2ints(BasisSet(mol, "sto-3g"), mol)
    assert len(ints) == 406
    assert np.max(np.abs(np.asarray(ints) - golden("h2o_sto3g")["crys"])) < 1e-12


def test_get2ints_h2o_631ss(golden):
    from pyquante_amd.Ints import get2ints
    from pyquante_amd.basis import BasisSet
    mol = _h2o_631ss()
    ints = get2ints(BasisSet(mol, "6-31g**"), mol)
    assert len(ints) == 52975
    ref = golden("h2o_631ss")["crys"]
    assert np.max(np.abs(np.asarray(ints) - ref)) < 1e-12


def test_direct_and_stored_jk_vs_oracle(golden):
    from pyquante_amd import Ints
    from pyquante_amd.basis import BasisSet
    mol = _h2o_631ss()
    bs = BasisSet(mol, "6-31g**")
    ref_ints = golden("h2o_631ss")["crys"]
    n = len(bs)
    rng = np.random.default_rng(4)
    D = rng.standard_normal((n, n)) * 0.1
    D = D + D.T
    Jo, Ko, Go = O.getJK(ref_ints, D)
    old = Ints.schwarz_cut
    Ints.schwarz_cut = 0.0
    try:
        S, h, I = Ints.getints(bs, mol)
        assert np.max(np.abs(Ints.getJ(I, D) - Jo)) < 1e-10
        assert np.max(np.abs(Ints.getK(I, D) - Ko)) < 1e-10
        assert np.max(np.abs(Ints.get2JmK(I, D) - Go)) < 1e-10
    finally:
        Ints.schwarz_cut = old
    # a plain flat array in the reference layout is digested on the GPU
    assert np.max(np.abs(Ints.getJ(ref_ints, D) - Jo)) < 1e-10
    assert np.max(np.abs(Ints.getK(list(ref_ints), D) - Ko)) < 1e-10


def test_get1ints_vs_oracle(golden):
    from pyquante_amd.Ints import get1ints
    from pyquante_amd.basis import BasisSet
    from pyquante_amd.engine import Engine
    from pyquante_amd.molecule import h2o_sto3g_config
    for mol, name in ((h2o_sto3g_config(), "sto-3g"), (_h2o_631ss(), "6-31g**")):
        bs = BasisSet(mol, name)
        # the drop-in default: V with the reference's Boys function
        S, h = get1ints(bs, mol)
        So, ho = O.get1ints(bs.bfs, mol.atoms, compat_nr=1)
        assert np.max(np.abs(S - So)) < 1e-12
        assert np.max(np.abs(h - ho)) < 1e-12
        # the engine's accurate V (the one the gradient differentiates)
        Se, Te, Ve = Engine(bs, mol).get1ints(compat_nr=0)
        _, he = O.get1ints(bs.bfs, mol.atoms, compat_nr=0)
        assert np.max(np.abs((Te + Ve).cpu().numpy() - he)) < 1e-11
        # the two differ by the reference's Boys error only
        assert 1e-9 < np.max(np.abs(he - ho)) < 1e-6
    g = golden("h2o_sto3g")
    mol = h2o_sto3g_config()
    S, h = get1ints(BasisSet(mol, "sto-3g"), mol)
    assert np.max(np.abs(S - g["S"])) < 1e-12
    # h is the compiled reference's T + V (cints.c, NR Boys EPS = 3e-7)
    assert np.max(np.abs(h - (g["T"] + g["V"]))) < 1e-12


def test_single_primitive_norms_6_31gss():
    """S_ii = 1 for every single-primitive function of 6-31G**, d included:
    the primitive norm of primitive-gto.c:76-83 normalises each Cartesian
    component, (2l-1)!!(2m-1)!!(2n-1)!! per component (xx and xy differ)."""
    from pyquante_amd.basis import BasisSet
    from pyquante_amd.engine import Engine
    mol = _h2o_631ss()
    bs = BasisSet(mol, "6-31g**")
    S = Engine(bs, mol).get1ints()[0].cpu().numpy()
    single = [i for i, f in enumerate(bs.bfs) if len(f.pexps) == 1]
    Ls = {sum(bs.bfs[i].powers) for i in single}
    assert Ls == {0, 1, 2}
    d = np.diag(S)[single]
    assert np.max(np.abs(d - 1.0)) < 1e-14


def test_rhf_h2o_sto3g_vs_oracle():
    """Configuration C1: the drop-in rhf (V with the reference's Boys
    function) against the oracle SCF over the reference's V, to 1e-9 Eh; and
    the accurate-V variant against its own oracle."""
    from oracle import scf
    from pyquante_amd.basis import BasisSet
    from pyquante_amd.hartree_fock import rhf
    from pyquante_amd.molecule import h2o_sto3g_config
    mol = h2o_sto3g_config()
    bs = BasisSet(mol, "sto-3g")
    e_ref = scf.rhf(bs.bfs, mol, ConvCriteria=1e-10, compat_nr=1)[0]
    e = rhf(mol, basis_data="sto-3g", ConvCriteria=1e-10, MaxIter=60)[0]
    print("C1 RHF %.12f, oracle %.12f" % (e, e_ref), flush=True)
    assert abs(e - e_ref) < 1e-9
    e_acc = scf.rhf(bs.bfs, mol, ConvCriteria=1e-10, compat_nr=0)[0]
    e2 = rhf(mol, basis_data="sto-3g", ConvCriteria=1e-10, MaxIter=60, reference_boys=False)[0]
    assert abs(e2 - e_acc) < 1e-9
    assert abs(e_acc - e_ref) > 1e-8  # the two V really differ


def test_rhf_h2o_631ss_reference_energy():
    from pyquante_amd.hartree_fock import rhf
    # Tests/h2o.py:16-23 (rhf defaults, 6-31G**): -76.011751 +- 1e-4
    e = rhf(_h2o_631ss(), ConvCriteria=1e-8, MaxIter=60)[0]
    assert abs(e - (-76.011751)) < 1e-4
    # Tests/WaterSweet.py:16 (DIIS): -76.011755864850628 to 4 places
    e2 = rhf(_h2o_631ss(), ConvCriteria=1e-8, MaxIter=60, DoAveraging=True)[0]
    assert abs(e2 - (-76.011755864850628)) < 1e-4


def test_water_dimer_direct_vs_stored_vs_oracle():
    from pyquante_amd.basis import BasisSet
    from pyquante_amd.engine import Engine, digest_stored
    from pyquante_amd.molecule import water_cluster
    mol = water_cluster((2, 1, 1), seed=1)
    bs = BasisSet(mol, "6-31g**")
    eng = Engine(bs, mol)
    n = eng.nbf
    rng = np.random.default_rng(5)
    D = rng.standard_normal((n, n)) * 0.05
    D = D + D.T
    ints = eng.get2ints()
    Dt = torch.as_tensor(D, device=eng.device)
    Js, Ks = digest_stored(ints, Dt)
    eng.build(0.0)
    Jd, Kd = eng.fock_jk(Dt)
    assert torch.max(torch.abs(Js - Jd)).item() < 1e-11
    assert torch.max(torch.abs(Ks - Kd)).item() < 1e-11
    Jo, Ko, _ = O.getJK(ints.cpu().numpy(), D)
    assert np.max(np.abs(Jd.cpu().numpy() - Jo)) < 1e-10
    assert np.max(np.abs(Kd.cpu().numpy() - Ko)) < 1e-10
    # unscreened work lists cover every unique shell quartet exactly once
    st = eng.stats()
    npair = len(bs.shells) * (len(bs.shells) + 1) // 2
    assert st["quartets_fp64"] == npair * (npair + 1) // 2


def test_shim_modules_match_oracle():
    from pyquante_amd import chgp, cints, crys
    from pyquante_amd.basis import BasisSet
    bfs = BasisSet(_h2o_631ss(), "6-31g**").bfs
    rng = np.random.default_rng(6)
    for _ in range(20):
        q = [bfs[i] for i in rng.integers(0, len(bfs), 4)]
        ref = O.contr_coulomb(*q)
        args = []
        for f in q:
            args += [f.pexps, f.pcoefs, f.pnorms, f.origin, f.powers]
        assert abs(crys.contr_coulomb(*args) - ref) < 1e-12
        # cints: THO with the reference's Boys function
        ref_t = O.contr_coulomb(*q, method=O.THO, compat_nr=1)
        assert abs(cints.contr_coulomb(*args) - ref_t) < 1e-12
        # chgp reproduces the reference's HGP values, EPS = 3e-7 Boys included
        ref_h = O.contr_coulomb(*q, method=O.HGP, compat_nr=1)
        assert abs(chgp.contr_coulomb(*args) - ref_h) < 1e-12
    a = bfs[5]
    v = crys.coulomb_repulsion(a.origin, a.pnorms[0], a.powers, a.pexps[0], a.origin, a.pnorms[0], a.powers,
                               a.pexps[0], a.origin, a.pnorms[0], a.powers, a.pexps[0], a.origin,
                               a.pnorms[0], a.powers, a.pexps[0])
    z = O.prim_eri(O.RYS, a.origin, a.pnorms[0], a.powers, a.pexps[0], a.origin, a.pnorms[0], a.powers,
                   a.pexps[0], a.origin, a.pnorms[0], a.powers, a.pexps[0], a.origin, a.pnorms[0],
                   a.powers, a.pexps[0])
    assert abs(v - z) < 1e-12 * abs(z)
    with pytest.raises(SystemError):  # > 40 primitives in total (crys.c:1633)
        crys.contr_coulomb(*([[1.0] * 11, [1.0] * 11, [1.0] * 11, (0, 0, 0), (0, 0, 0)] * 4))
    with pytest.raises(ValueError):  # L_total > 9 has no Rys rule
        crys.coulomb_repulsion((0, 0, 0), 1, (3, 0, 0), 1, (0, 0, 0), 1, (3, 0, 0), 1, (0, 0, 0), 1, (2, 0, 0), 1,
                               (0, 0, 0), 1, (2, 0, 0), 1)


def test_mixed_precision_close_to_fp64():
    from pyquante_amd import _abi
    from pyquante_amd.basis import BasisSet
    from pyquante_amd.engine import Engine
    from pyquante_amd.molecule import water_cluster
    mol = water_cluster((2, 2, 1), seed=2)
    eng = Engine(BasisSet(mol, "6-31g**"), mol)
    n = eng.nbf
    rng = np.random.default_rng(7)
    C = rng.standard_normal((n, 20)) * 0.1
    D = torch.as_tensor(C @ C.T, device=eng.device)
    eng.build(1e-12, 0.0)
    J64, K64 = eng.fock_jk(D)
    eng.build(1e-12, 1e-4)
    st = eng.stats()
    assert st["quartets_fp32"] > 0
    J32, K32 = eng.fock_jk(D, precision=_abi.PQ_PREC_MIXED)
    e64 = float(torch.sum(D * (2 * J64 - K64)))
    e32 = float(torch.sum(D * (2 * J32 - K32)))
    assert abs(e64 - e32) < 1e-6
    assert torch.max(torch.abs(J64 - J32)).item() < 1e-6


def test_primitive_screening_bound():
    """Skipping primitive pairs (factor product below 1e-4 * cut) moves J/K
    by far less than the shell-level cut itself and leaves the quartet count
    alone; the unscreened build is pinned against the oracle elsewhere."""
    from pyquante_amd.basis import BasisSet
    from pyquante_amd.engine import Engine
    from pyquante_amd.molecule import water_cluster
    mol = water_cluster((2, 2, 1), seed=5)
    eng = Engine(BasisSet(mol, "6-31g**"), mol)
    n = eng.nbf
    rng = np.random.default_rng(9)
    C = rng.standard_normal((n, 20)) * 0.1
    D = torch.as_tensor(C @ C.T, device=eng.device)
    eng.build(0.0)
    Je, Ke = eng.fock_jk(D)
    eng.set_prim_screen(0.0)
    eng.build(1e-12)
    Ja, Ka = eng.fock_jk(D)
    sa = eng.stats()
    eng.set_prim_screen(1e-4)
    eng.build(1e-12)
    Jb, Kb = eng.fock_jk(D)
    sb = eng.stats()
    print("prim quartets %d -> %d" % (sa["prim_quartets_fp64"], sb["prim_quartets_fp64"]), flush=True)
    assert sb["quartets_fp64"] == sa["quartets_fp64"]
    assert sb["prim_quartets_fp64"] < 0.9 * sa["prim_quartets_fp64"]
    dprim = max(torch.max(torch.abs(Ja - Jb)).item(), torch.max(torch.abs(Ka - Kb)).item())
    dshell = max(torch.max(torch.abs(Ja - Je)).item(), torch.max(torch.abs(Ka - Ke)).item())
    print("max |d| from primitive screening %.2e, from the shell cut %.2e" % (dprim, dshell), flush=True)
    assert dprim < 1e-11
    assert dshell < 1e-8
    # every primitive again once switched off
    eng.set_prim_screen(0.0)
    eng.build(1e-12)
    assert eng.stats()["prim_quartets_fp64"] == sa["prim_quartets_fp64"]


def test_function_quartet_count_unscreened():
    from pyquante_amd.basis import BasisSet
    from pyquante_amd.engine import Engine
    from pyquante_amd.molecule import water_cluster
    mol = water_cluster((2, 1, 1), seed=3)
    bs = BasisSet(mol, "6-31g**")
    eng = Engine(bs, mol).build(0.0)
    n = eng.nbf
    npair = n * (n + 1) // 2
    assert eng.stats()["fquartets_fp64"] == npair * (npair + 1) // 2


def _contr_args(fs):
    args = []
    for f in fs:
        args += [f.pexps, f.pcoefs, f.pnorms, f.origin, f.powers]
    return args


def test_cints_dropin_vs_compiled_cints(golden):
    """The cints drop-in (THO, cints.c:80-153, with its EPS = 3e-7 Boys
    function) against the compiled reference's cints values: all 406
    H2O/STO-3G integrals, every 5th of the 52,975 H2O/6-31G** integrals and the
    random primitive classes up to L_total = 9."""
    from pyquante_amd import cints
    from pyquante_amd._compat import THO, contr_coulomb_batch
    from pyquante_amd.basis import BasisSet
    from pyquante_amd.molecule import h2o_sto3g_config

    def unique(bfs, stride=1):
        qs, n, c = [], len(bfs), 0
        for i in range(n):
            for j in range(i + 1):
                for k in range(n):
                    for l in range(k + 1):
                        if i * (i + 1) // 2 + j <= k * (k + 1) // 2 + l:
                            if c % stride == 0:
                                qs.append(tuple((f.pexps, f.pcoefs, f.pnorms, f.origin, f.powers)
                                                for f in (bfs[i], bfs[j], bfs[k], bfs[l])))
                            c += 1
        return qs

    g = golden("h2o_sto3g")
    bfs = BasisSet(h2o_sto3g_config(), "sto-3g").bfs
    qs = unique(bfs)
    v = contr_coulomb_batch(qs, THO, 1)
    idx = [O.ijkl2intindex(*t) for t in _unique_ijkl(len(bfs))]
    ref = g["cints"][idx]
    print("STO-3G cints max|d| %.2e (crys differs by %.2e)" % (np.max(np.abs(v - ref)),
                                                              np.max(np.abs(g["crys"] - g["cints"]))), flush=True)
    assert np.max(np.abs(v - ref)) < 1e-12
    g5 = golden("h2o_631ss")
    bfs = BasisSet(_h2o_631ss(), "6-31g**").bfs
    v5 = contr_coulomb_batch(unique(bfs, 5), THO, 1)
    assert len(v5) == len(g5["cints5"])
    print("6-31G** cints (10,595) max|d| %.2e" % np.max(np.abs(v5 - g5["cints5"])), flush=True)
    assert np.max(np.abs(v5 - g5["cints5"])) < 1e-12
    pc = golden("prim_classes")
    vals = []
    for t in range(len(pc["powers"])):
        pw = pc["powers"][t].reshape(4, 3)
        o = pc["origins"][t].reshape(4, 3)
        a = pc["alphas"][t]
        args = []
        for f in range(4):
            args += [tuple(o[f]), 1.0, tuple(int(x) for x in pw[f]), float(a[f])]
        vals.append(cints.coulomb_repulsion(*args))
    vals = np.array(vals)
    err = np.max(np.abs(vals - pc["cints"]) / np.maximum(1.0, np.abs(pc["cints"])))
    print("primitive classes (L <= 9) cints max rel|d| %.2e" % err, flush=True)
    assert err < 1e-12
    # with full-precision Boys values the THO path gives the Rys (crys) value
    cints.accurate_boys = True
    try:
        acc = cints.contr_coulomb(*_contr_args(bfs[3:7]))
    finally:
        cints.accurate_boys = False
    assert abs(acc - O.contr_coulomb(*bfs[3:7])) < 1e-12


def _unique_ijkl(n):
    out = []
    for i in range(n):
        for j in range(i + 1):
            for k in range(n):
                for l in range(k + 1):
                    if i * (i + 1) // 2 + j <= k * (k + 1) // 2 + l:
                        out.append((i, j, k, l))
    return out


def test_getj_getk_of_one_density_share_a_build():
    """getJ and getK of one density (the reference's UHF loop calls them
    separately, hartree_fock.py:233-236) cost one direct build: Integrals
    keeps its last two builds and answers a density equal to a kept one
    from them.  Results equal a fresh get2JmK; a changed density, or one
    edited in place, is rebuilt; returned matrices are copies."""
    from pyquante_amd.Ints import getints, get2JmK, getJ, getK
    from pyquante_amd.basis import BasisSet
    mol = _h2o_631ss()
    bs = BasisSet(mol, "6-31g**")
    _, _, Ints = getints(bs, mol)
    calls = []
    real = Ints.engine.fock_jk

    def counting(*a, **k):
        calls.append(1)
        return real(*a, **k)

    Ints.engine.fock_jk = counting
    rng = np.random.default_rng(3)
    n = len(bs.bfs)
    A = rng.standard_normal((n, 6))
    Da = A @ A.T
    B = rng.standard_normal((n, 6))
    Db = B @ B.T
    Ja, Jb = getJ(Ints, Da), getJ(Ints, Db)
    Ka, Kb = getK(Ints, Da), getK(Ints, Db)
    assert len(calls) == 2
    Ga = get2JmK(Ints, Da)
    assert len(calls) == 2 and np.max(np.abs(Ga - (2 * Ja - Ka))) < 1e-12
    Jt = getJ(Ints, torch.as_tensor(Da, device="cuda"))  # a device density: a device J
    assert len(calls) == 2 and np.max(np.abs(Jt.cpu().numpy() - Ja)) < 1e-12
    Jt += 1.0  # a copy: the kept J is untouched
    assert np.max(np.abs(getJ(Ints, Da) - Ja)) < 1e-12
    Da[1, 2] += 0.01
    Da[2, 1] += 0.01
    J2, K2 = getJ(Ints, Da), getK(Ints, Da)
    assert len(calls) == 3
    Ints.engine.fock_jk = real
    Jf, Kf = real(torch.as_tensor(Da, device="cuda"))
    assert np.max(np.abs(J2 - Jf.cpu().numpy())) < 1e-12 and np.max(np.abs(K2 - Kf.cpu().numpy())) < 1e-12
