"""GPU Head-Gordon-Pople ERIs (the chgp drop-in) against the compiled
reference chgp.c and against the Rys engine.

- compat mode (the reference's EPS = 3e-7 Boys function) against the
  compiled chgp golden values: 1e-12 absolute;
- full-precision Boys against the Rys engine and the compiled crys.c values:
  1e-12 absolute (the methods agree to rounding once both use exact Boys
  values);
- vrr leaves against the (ss|ss) closed form pref * F_m(T);
- the shell-quartet kernel (VRR per primitive quartet, HRR once per
  contracted block, Boys values from Rys moments) against the Rys store
  kernels and the per-function kernel on every class.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _unique_quartets(n):
    out = []
    for i in range(n):
        for j in range(i + 1):
            for k in range(n):
                for l in range(k + 1):
                    if i * (i + 1) // 2 + j >= k * (k + 1) // 2 + l:
                        out.append((i, j, k, l))
    return out


def _args(bfs, q):
    return tuple((bfs[f].pexps, bfs[f].pcoefs, bfs[f].pnorms, bfs[f].origin, bfs[f].powers) for f in q)


def _batch(bfs, quartets, compat_nr):
    from pyquante_amd._compat import HGP, contr_coulomb_batch
    return contr_coulomb_batch([_args(bfs, q) for q in quartets], HGP, compat_nr)


def test_hgp_compat_h2o_sto3g_vs_compiled_chgp(golden):
    from pyquante_amd.basis import BasisSet
    from pyquante_amd.molecule import h2o_sto3g_config
    mol = h2o_sto3g_config()
    bfs = BasisSet(mol, "sto-3g").bfs
    qs = _unique_quartets(len(bfs))
    idx = np.array([O.ijkl2intindex(*q) for q in qs])
    v = _batch(bfs, qs, 1)
    ref = golden("h2o_sto3g")["chgp"][idx]
    assert np.max(np.abs(v - ref)) < 1e-12
    # the same quartets with exact Boys values agree with crys.c instead
    v2 = _batch(bfs, qs, 0)
    assert np.max(np.abs(v2 - golden("h2o_sto3g")["crys"][idx])) < 1e-12


def test_hgp_compat_h2o_631ss_vs_compiled_chgp(golden):
    from pyquante_amd.basis import BasisSet
    from pyquante_amd.molecule import Molecule
    r = 1 / 0.52918
    mol = Molecule("h2o", [(8, (0, 0, 0)), (1, (r, 0, 0)), (1, (0, r, 0))])
    bfs = BasisSet(mol, "6-31g**").bfs
    g = golden("h2o_631ss")
    want = {int(i): t for t, i in enumerate(g["index5"])}
    qs, rows = [], []
    for q in _unique_quartets(len(bfs)):
        t = want.get(O.ijkl2intindex(*q))
        if t is not None:
            qs.append(q)
            rows.append(t)
    assert len(qs) == len(g["index5"])
    v = _batch(bfs, qs, 1)
    print("hgp compat 6-31G** max |d| %.2e" % np.max(np.abs(v - g["chgp5"][rows])), flush=True)
    assert np.max(np.abs(v - g["chgp5"][rows])) < 1e-12
    v2 = _batch(bfs, qs, 0)
    assert np.max(np.abs(v2 - g["crys"][g["index5"][rows]])) < 1e-12


def test_hgp_primitive_classes_vs_compiled_chgp(golden):
    from pyquante_amd._compat import HGP, contr_coulomb_batch
    g = golden("prim_classes")
    P, Org, A = g["powers"], g["origins"], g["alphas"]
    keep = []
    for t in range(len(P)):
        p = P[t].reshape(4, 3)
        box_e = np.prod(p[0] + p[1] + 1)
        box_f = np.prod(p[2] + p[3] + 1)
        if box_e <= 12 and box_f <= 12 and p.sum() <= 8:
            keep.append(t)
    assert len(keep) > 100
    qs = [tuple(([A[t, q]], [1.0], [1.0], Org[t, 3 * q:3 * q + 3], P[t, 3 * q:3 * q + 3]) for q in range(4))
          for t in keep]
    v = contr_coulomb_batch(qs, HGP, 1)
    assert np.allclose(v, g["chgp"][keep], rtol=1e-11, atol=1e-14)
    v0 = contr_coulomb_batch(qs, HGP, 0)
    assert np.allclose(v0, g["crys"][keep], rtol=1e-11, atol=1e-14)
    # beyond the supported leaf size: a clear error, not garbage
    big = (([1.0], [1.0], [1.0], (0, 0, 0), (2, 1, 1)),) * 4
    with pytest.raises(ValueError):
        contr_coulomb_batch([big], HGP, 1)


def test_chgp_module_and_vrr_closed_form():
    from pyquante_amd import chgp
    A, B, C, D = (0.0, 0.0, 0.0), (0.3, -0.2, 0.5), (1.1, 0.4, -0.3), (0.9, 1.2, 0.2)
    aa, ab, ac, ad = 1.3, 0.7, 2.1, 0.45
    for m in range(0, 6):
        v = chgp.vrr(A, 1.0, (0, 0, 0), aa, B, 1.0, ab, C, 1.0, (0, 0, 0), ac, D, 1.0, ad, m)
        zeta, eta = aa + ab, ac + ad
        P = [(aa * A[i] + ab * B[i]) / zeta for i in range(3)]
        Q = [(ac * C[i] + ad * D[i]) / eta for i in range(3)]
        rab2 = sum((A[i] - B[i]) ** 2 for i in range(3))
        rcd2 = sum((C[i] - D[i]) ** 2 for i in range(3))
        T = zeta * eta / (zeta + eta) * sum((P[i] - Q[i]) ** 2 for i in range(3))
        k = math.sqrt(2) * math.pi ** 1.25
        pref = k / zeta * math.exp(-aa * ab / zeta * rab2) * k / eta * math.exp(-ac * ad / eta * rcd2)
        # the reference starts its downward recursion at order m with its NR Fgamma
        F = [0.0] * (m + 1)
        F[m] = O.fgamma_nr(m, T)
        for j in range(m - 1, -1, -1):
            F[j] = (2 * T * F[j + 1] + math.exp(-T)) / (2 * j + 1)
        assert abs(v - pref / math.sqrt(zeta + eta) * F[m]) < 1e-13 * abs(v)
    # vrr with m = 0 is the primitive (ps|ds) integral of coulomb_repulsion
    v = chgp.vrr(A, 0.7, (1, 0, 0), aa, B, 1.1, ab, C, 0.9, (0, 1, 1), ac, D, 1.3, ad, 0)
    w = chgp.coulomb_repulsion(A, 0.7, (1, 0, 0), aa, B, 1.1, (0, 0, 0), ab, C, 0.9, (0, 1, 1), ac, D, 1.3,
                               (0, 0, 0), ad)
    assert abs(v - w) < 1e-14 * abs(w)
    z = O.prim_eri(O.HGP, A, 0.7, (1, 0, 0), aa, B, 1.1, (0, 0, 0), ab, C, 0.9, (0, 1, 1), ac, D, 1.3,
                   (0, 0, 0), ad, compat_nr=1)
    assert abs(w - z) < 1e-12 * abs(z)


def test_get2ints_hgp_h2o_631ss(golden):
    from pyquante_amd.basis import BasisSet
    from pyquante_amd.engine import Engine
    from pyquante_amd.molecule import Molecule
    r = 1 / 0.52918
    mol = Molecule("h2o", [(8, (0, 0, 0)), (1, (r, 0, 0)), (1, (0, r, 0))])
    eng = Engine(BasisSet(mol, "6-31g**"), mol)
    g = golden("h2o_631ss")
    exact = eng.get2ints("hgp", 0).cpu().numpy()
    assert len(exact) == 52975
    assert np.max(np.abs(exact - g["crys"])) < 1e-12
    compat = eng.get2ints("hgp", 1).cpu().numpy()
    assert np.max(np.abs(compat[g["index5"]] - g["chgp5"])) < 1e-12


def test_get2ints_hgp_shell_quartets_every_class():
    """The shell-quartet HGP kernel (get2ints "hgp", full-precision Boys) over
    a water dimer in 6-31G**, which holds all 21 classes up to (dd|dd):
    against the Rys store kernels over the whole array, and against the
    per-function HGP kernel (independent VRR per leaf, series Boys values) on
    a seeded sample of quartets from every class."""
    import torch

    from pyquante_amd.basis import BasisSet
    from pyquante_amd.engine import Engine
    from pyquante_amd.molecule import water_cluster
    mol = water_cluster((2, 1, 1), seed=4)
    bs = BasisSet(mol, "6-31g**")
    eng = Engine(bs, mol)
    hgp = eng.get2ints("hgp", 0)
    rys = eng.get2ints("rys")
    assert hgp.shape == rys.shape
    assert torch.isfinite(hgp).all()
    assert float((hgp - rys).abs().max()) < 1e-12
    bfs = bs.bfs
    L = [sum(f.powers) for f in bfs]
    rng = np.random.default_rng(11)
    by_class = {}
    n = len(bfs)
    for _ in range(60000):
        i, j, k, l = (int(x) for x in rng.integers(0, n, 4))
        if i < j:
            i, j = j, i
        if k < l:
            k, l = l, k
        if i * (i + 1) // 2 + j < k * (k + 1) // 2 + l:
            i, j, k, l = k, l, i, j
        key = tuple(sorted((L[i], L[j]), reverse=True)) + tuple(sorted((L[k], L[l]), reverse=True))
        by_class.setdefault(key, [])
        if len(by_class[key]) < 40:
            by_class[key].append((i, j, k, l))
    qs = [q for v in by_class.values() for q in v]
    assert max(L) == 2 and (2, 2, 2, 2) in by_class
    per_function = _batch(bfs, qs, 0)
    idx = [i * (i + 1) // 2 + j for i, j, _, _ in qs], [k * (k + 1) // 2 + l for _, _, k, l in qs]
    flat = [ij * (ij + 1) // 2 + kl for ij, kl in zip(*idx)]
    shell = hgp.cpu().numpy()[flat]
    assert np.max(np.abs(shell - np.asarray(per_function))) < 1e-12


def test_rhf_hgp_backend_matches_rys():
    from pyquante_amd import Ints
    from pyquante_amd.hartree_fock import rhf
    from pyquante_amd.molecule import water_cluster
    mol = water_cluster((2, 1, 1), seed=4)
    e_rys, _, _ = rhf(mol, basis_data="6-31g**", ConvCriteria=1e-10, MaxIter=60, DoAveraging=True)
    old = Ints.hgp_compat_boys
    try:
        Ints.hgp_compat_boys = False
        e_hgp, _, _ = rhf(mol, basis_data="6-31g**", ConvCriteria=1e-10, MaxIter=60, DoAveraging=True,
                          method="hgp")
        Ints.hgp_compat_boys = True
        e_hgp_nr, _, _ = rhf(mol, basis_data="6-31g**", ConvCriteria=1e-10, MaxIter=60, DoAveraging=True,
                             method="hgp")
    finally:
        Ints.hgp_compat_boys = old
    print("rhf water dimer: rys %.12f hgp %.12f hgp(NR Boys) %.12f" % (e_rys, e_hgp, e_hgp_nr), flush=True)
    assert abs(e_hgp - e_rys) < 1e-9
    # the reference's EPS = 3e-7 Boys function moves the energy by far less than 1e-4 Eh
    assert abs(e_hgp_nr - e_rys) < 1e-4
