"""Host-side logic and the C ABI surface (no GPU compute)."""
import ctypes
import os
import re

import numpy as np
import pytest

from oracle import oracle as O
from pyquante_amd.basis import BasisSet, sym2powerlist
from pyquante_amd.molecule import h2o_sto3g_config, water_cluster

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "pq_abi.h")).read()
    names = set(re.findall(r"\b(pq_[a-z0-9_]+)\s*\(", hdr))
    assert len(names) >= 15
    lib = ctypes.CDLL(os.path.join(ROOT, "pyquante_amd", "libpqhip.so"))
    for n in sorted(names):
        assert hasattr(lib, n), n


def test_ijkl2intindex_abi_bit_exact():
    from pyquante_amd.Ints import ijkl2intindex
    rng = np.random.default_rng(2)
    for q in rng.integers(0, 420, size=(3000, 4)):
        assert ijkl2intindex(*map(int, q)) == O.ijkl2intindex(*map(int, q))
    # past the reference's 32-bit overflow (nbf = 400)
    assert ijkl2intindex(399, 398, 399, 399) == O.ijkl2intindex(399, 398, 399, 399)


def test_basis_order_and_counts():
    bs = BasisSet(h2o_sto3g_config(), "sto-3g")
    assert len(bs) == 7
    assert [f.powers for f in bs.bfs[2:5]] == sym2powerlist["P"]
    w = water_cluster((2, 2, 2), seed=0)
    b2 = BasisSet(w, "6-31g**")
    assert len(b2) == 200 and len(b2.shells) == 96
    d = [f.powers for f in b2.bfs[0:25] if sum(f.powers) == 2]
    assert d == sym2powerlist["D"]


def test_water_cluster_seeded_and_rigid():
    a = water_cluster((2, 2, 4), seed=5).coords()
    b = water_cluster((2, 2, 4), seed=5).coords()
    assert np.array_equal(a, b) and a.shape == (48, 3)
    r = np.linalg.norm(a[1] - a[0]) * 0.529177249
    assert abs(r - 0.9572) < 1e-12
    v1, v2 = a[1] - a[0], a[2] - a[0]
    ang = np.degrees(np.arccos(v1 @ v2 / np.linalg.norm(v1) / np.linalg.norm(v2)))
    assert abs(ang - 104.52) < 1e-9


def test_gpu_paths_fail_loudly_without_device():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    from pyquante_amd import _abi
    from pyquante_amd.engine import Engine
    with pytest.raises(_abi.PQError):
        Engine(BasisSet(h2o_sto3g_config(), "sto-3g"), h2o_sto3g_config())
    from pyquante_amd import crys
    with pytest.raises(_abi.PQError):
        crys.Fgamma(0, 1.0)


def _parse_table32():
    path = os.path.join(os.path.dirname(__file__), "..", "pyquante_amd", "csrc", "rys_table32.h")
    src = open(path).read()
    off = [int(v) for v in re.search(r"rys_off32\[\d+\] = \{([^}]*)\}", src).group(1).split(",")]
    xa = [float(v) for v in re.search(r"rys_xasym32\[\d+\] = \{([^}]*)\}", src).group(1).split(",")]
    body = re.search(r"rys_coef32\[\d+\] = \{(.*?)\};", src, re.S).group(1)
    coef = np.array([float(v.strip().rstrip("f")) for v in body.split(",")], dtype=np.float32)
    stride = int(re.search(r"#define RYS_STRIDE32 (\d+)", src).group(1))
    deg = int(re.search(r"#define RYS_DEG32 (\d+)", src).group(1))
    return off, xa, coef, stride, deg


def test_fp32_rys_table_moments():
    """The FP32 root table, evaluated in float32 Horner as the device does,
    reproduces the Boys moments sum_i w_i u_i^k = F_k(X), k < 2n, to float
    accuracy; the FP64 large-X limit takes over at rys_xasym32."""
    off, xa, coef, stride, deg = _parse_table32()
    rng = np.random.default_rng(3)
    for n in range(1, 6):
        X = np.concatenate([[0.0, 0.5, xa[n] - 1e-9], rng.uniform(0, xa[n], 60), [xa[n], xa[n] + 3.7]])
        for x in X:
            if x < xa[n]:
                k = int(x)
                s = np.float32(2.0 * (x - k) - 1.0)
                base = off[n - 1] + k * 2 * n * stride
                vals = []
                for f in range(2 * n):
                    c = coef[base + f * stride: base + f * stride + deg + 1]
                    v = c[deg]
                    for j in range(deg - 1, -1, -1):
                        v = np.float32(v * s + c[j])
                    vals.append(float(v))
                u, w = np.array(vals[:n]), np.array(vals[n:])
            else:
                u, w = O.rys_roots(n, x)  # the large-X branch is FP64 on the device too
            for m in range(2 * n):
                ref = O.boys(m, x)
                got = float(np.sum(w * u ** m))
                assert abs(got - ref) <= 1e-6 * ref + 1e-30, (n, x, m, got, ref)


def test_stale_fock_graph_raises_on_host():
    """A captured Fock build refuses to replay once its engine's work-list
    generation has moved (the check runs on the host, before any replay)."""
    from pyquante_amd.engine import FockGraph, StaleGraphError

    class _Eng(object):
        gen = 7

        def generation(self):
            return self.gen

    class _Graph(object):
        replays = 0

        def replay(self):
            _Graph.replays += 1

    fg = FockGraph.__new__(FockGraph)
    fg.engine = _Eng()
    fg.graph = _Graph()
    fg.generation = 7
    fg.J = fg.K = None
    fg()
    assert _Graph.replays == 1
    fg.engine.gen = 8  # e.g. build() or set_group_caps() replaced the work lists
    with pytest.raises(StaleGraphError):
        fg()
    assert _Graph.replays == 1


def _boys_table():
    txt = open(os.path.join(ROOT, "pyquante_amd", "csrc", "boys_table.h")).read()
    macros = dict((k, int(v)) for k, v in re.findall(r"#define (BOYS_\w+) (\d+)", txt))
    body = txt[txt.index("boys_grid["):]
    body = body[body.index("{") + 1:body.index("};")]
    vals = np.array([float(x) for x in body.replace("\n", " ").split(",") if x.strip()])
    return macros, vals.reshape(macros["BOYS_NG"], macros["BOYS_STRIDE"])


def test_boys_grid_table_against_oracle():
    """The grid table of the J/K kernels (F_m at T = g/4) and the device's
    evaluation scheme (Taylor series of order 10 about the nearest grid
    point, asymptotic form above T = 48), restated in numpy, reproduce the
    oracle's accurate Boys function to 2e-15 relative for m <= 3."""
    mc, tab = _boys_table()
    inv_h, K, tasym = mc["BOYS_INV_H"], mc["BOYS_K"], mc["BOYS_TASYM"]
    # the grid values themselves
    for g in range(0, mc["BOYS_NG"], 7):
        for m in range(mc["BOYS_MTAB"]):
            ref = O.boys(m, g / inv_h)
            assert abs(tab[g, m] - ref) <= 4e-16 * ref, (g, m)
    rng = np.random.default_rng(4)
    Ts = np.concatenate([rng.uniform(0, 60, 400), [0.0, 1e-12, 0.124999, 0.125, 47.99, 48.0, 48.01, 200.0]])
    for T in Ts:
        for m in range(4):
            if T < tasym:
                g = int(T * inv_h + 0.5)
                x = g / inv_h - T
                v = tab[g, m + K]
                for k in range(K - 1, -1, -1):
                    v = v * (x / (k + 1)) + tab[g, m + k]
            else:
                v = 0.88622692545275801365 / np.sqrt(T)
                for j in range(1, m + 1):
                    v *= (j - 0.5) / T
            ref = O.boys(m, T)
            assert abs(v - ref) <= 2e-15 * ref, (T, m, v, ref)


def test_cgbf_list_norms_are_checked():
    """A flat CGBF list is regrouped into shells; the engine derives every
    normalisation itself, so CGBFs carrying other norms are refused rather
    than silently integrated with the engine's (CGBF.py:34,
    primitive-gto.c:76-83)."""
    import copy
    from pyquante_amd.engine import shells_of
    bs = BasisSet(water_cluster((1, 1, 1), seed=0), "6-31g**")
    flat = list(bs.bfs)
    sh = shells_of(flat)
    assert [s[0] for s in sh] == [s.L for s in bs.shells]
    bad = copy.deepcopy(flat)
    bad[5].norm = 0.5
    with pytest.raises(ValueError, match="norm"):
        shells_of(bad)
    bad = copy.deepcopy(flat)
    bad[2].pnorms[0] *= 1.01
    with pytest.raises(ValueError, match="primitive norm"):
        shells_of(bad)


def test_wavefunction_density_matrices():
    """Densities of the force module's wavefunctions (Wavefunction.py:143-199)
    on random orbitals: the fixed-occupation ones with integer occupations
    equal the unrestricted ones, with the alpha orbital energies weighting
    the beta energy-weighted density as the reference's mk_auger_Qmatrix
    does; fractional occupations scale each orbital's outer product; and the
    closed-shell density is the unrestricted alpha one."""
    from pyquante_amd.force import FixedOccWavefunction, UHFWavefunction, Wavefunction
    rng = np.random.default_rng(7)
    n = 9
    ca, cb = rng.standard_normal((n, n)), rng.standard_normal((n, n))
    ea, eb = np.sort(rng.standard_normal(n)), np.sort(rng.standard_normal(n))
    fx = FixedOccWavefunction(ca, cb, ea, eb, [1.0] * 4, [1.0] * 3)
    u = UHFWavefunction((ca, cb), (ea, ea), 4, 3)
    for x, y in zip(fx.mk_auger_dens() + fx.mk_auger_Qmatrix(), u.mkdens() + u.mkQmatrix()):
        assert np.max(np.abs(x - y)) < 1e-13
    occ = [1.0, 0.5, 0.25]
    Da, _ = FixedOccWavefunction(ca, cb, ea, eb, occ, occ).mk_auger_dens()
    ref = sum(o * np.outer(ca[:, i], ca[:, i]) for i, o in enumerate(occ))
    assert np.max(np.abs(Da - ref)) < 1e-13
    Qa, Qb = FixedOccWavefunction(ca, cb, ea, eb, occ, occ).mk_auger_Qmatrix()
    refb = sum(ea[i] * o * np.outer(cb[:, i], cb[:, i]) for i, o in enumerate(occ))
    assert np.max(np.abs(Qb - refb)) < 1e-13
    w = Wavefunction(ca, ea, 4)
    assert np.max(np.abs(w.mkdens() - u.mkdens()[0])) < 1e-13
    assert np.max(np.abs(w.mkQmatrix() - u.mkQmatrix()[0])) < 1e-13


def test_diis_incremental_gram_matches_full_system():
    """hartree_fock.DIIS keeps the Gram matrix of the errors one row per call
    and solves on the host; on CPU tensors its extrapolated Fock matrices
    and maxerr must equal the full (k+1)-system of Convergence.py:88-136
    rebuilt from every stored error at each call."""
    import torch
    from pyquante_amd.hartree_fock import DIIS
    rng = np.random.default_rng(1)
    n = 6
    sym = lambda a: a + a.T  # noqa: E731
    S = torch.as_tensor(np.eye(n) + 0.1 * sym(rng.standard_normal((n, n))))
    d = DIIS(S)
    Fs, Es = [], []
    for it in range(12):  # past the first buffer growth (8)
        F = torch.as_tensor(sym(rng.standard_normal((n, n))))
        C = rng.standard_normal((n, 3))
        D = torch.as_tensor(C @ C.T)
        out = d.getF(F, D)
        e = (F @ D @ S - S @ D @ F).reshape(-1)
        Fs.append(F)
        Es.append(e)
        k = len(Es)
        a = np.zeros((k + 1, k + 1))
        for i in range(k):
            for j in range(k):
                a[i, j] = float(Es[i] @ Es[j])
        a[k, :k] = a[:k, k] = -1.0
        b = np.zeros(k + 1)
        b[k] = -1.0
        c = np.linalg.solve(a, b)
        ref = sum(c[i] * Fs[i] for i in range(k))
        assert float((out - ref).abs().max()) < 1e-10, it
        assert abs(d.maxerr - float(e.abs().max())) < 1e-14


def test_batched_diis_matches_per_molecule_systems():
    """batch._batched_diis_fock (Gram matrix one row per call, the last 8
    entries, host solve) on CPU tensors: each molecule's extrapolated Fock
    matrix equals the full DIIS system of its own last 8 errors."""
    import torch
    from pyquante_amd.batch import _batched_diis_fock, _DiisState
    rng = np.random.default_rng(3)
    nm, n = 5, 4
    st = _DiisState()
    hist = []
    for it in range(11):  # past the 8-entry window
        F = torch.as_tensor(rng.standard_normal((nm, n, n)))
        err = torch.as_tensor(rng.standard_normal((nm, n, n)))
        out = _batched_diis_fock(st, F, err)
        hist = (hist + [(F, err)])[-8:]
        k = len(hist)
        if k < 2:
            assert torch.equal(out, F)
            continue
        for m in range(nm):
            E = np.array([h[1][m].reshape(-1).numpy() for h in hist])
            B = -np.ones((k + 1, k + 1))
            B[:k, :k] = E @ E.T
            B[k, k] = 0.0
            r = np.zeros(k + 1)
            r[k] = -1.0
            c = np.linalg.solve(B, r)[:k]
            ref = sum(c[j] * hist[j][0][m].numpy() for j in range(k))
            assert np.abs(out[m].numpy() - ref).max() < 1e-9, (it, m)
