"""The host work-list builder (pyquante_amd/csrc/pq_worklist.h) on the CPU,
built with AddressSanitizer and UndefinedBehaviorSanitizer
(tests/native/worklist_check.cpp): for water clusters in 6-31G** -- one
molecule, and a batch of molecules as the batched engine groups them -- with
Schwarz bounds and primitive screening of realistic shape, every class's
tiles must cover each kept (bra, ket) shell quartet exactly once, whole
(primitive ranges partitioning the bra's contraction), in the bucket its
bound selects, with the screened quartets absent, and the bra runs must
partition the tiles, in order, by bra; any out-of-bounds access or
undefined behaviour fails the run."""
import math
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    exe = str(tmp_path_factory.mktemp("wl") / "worklist_check")
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
                    "-fno-sanitize-recover=undefined", "-o", exe, os.path.join(HERE, "native", "worklist_check.cpp")],
                   check=True)
    return exe


def _pairs(molecules, basis="6-31g**"):
    """Shell pairs within each molecule as the engine forms them (A >= B by
    shell index, stored with l_A >= l_B): pair class, shells, unscreened
    contraction length, a Schwarz-like bound Q and per-primitive factors
    (s-type overlap magnitudes: the shape, not the value, matters here)."""
    from pyquante_amd.basis import BasisSet
    shells, sh_mol = [], []
    for m, mol in enumerate(molecules):
        for sh in BasisSet(mol, basis).shells:
            shells.append((sh.L, np.array(sh.exps), np.array(sh.coefs), np.array(sh.origin, dtype=float)))
            sh_mol.append(m)
    pairs = []
    for A in range(len(shells)):
        for B in range(A + 1):
            if sh_mol[A] != sh_mol[B]:
                continue
            s1, s2 = (A, B) if shells[A][0] >= shells[B][0] else (B, A)
            l1, a1, c1, r1 = shells[s1]
            l2, a2, c2, r2 = shells[s2]
            R2 = float(np.sum((r1 - r2) ** 2))
            a, b = np.meshgrid(a1, a2, indexing="ij")
            c = np.abs(np.outer(c1, c2))
            p = a + b
            f = np.sort((c * np.exp(-a * b / p * R2) * (math.pi / p) ** 1.5 * p ** 0.25).ravel())[::-1]
            cls = l1 * (l1 + 1) // 2 + l2
            pairs.append((cls, s1, s2, f, float(f.sum())))
    return pairs, sh_mol


def _write(path, pairs, sh_mol, nmol, cut, fp32cut, prim_rel, split, caps):
    fmax = max(p[3][0] for p in pairs)
    tau = cut * prim_rel if cut > 0 else 0.0
    pp_s = []
    for cls, s1, s2, f, q in pairs:
        n = 1
        while tau > 0 and n < len(f) and f[n] * fmax >= tau:
            n += 1
        pp_s.append(n if tau > 0 else len(f))
    order = []
    for c in range(6):
        ids = [i for i, p in enumerate(pairs) if p[0] == c]
        ids.sort(key=lambda i: (sh_mol[pairs[i][1]], -len(pairs[i][3]), -pairs[i][4]))
        order.append(ids)
    with open(path, "wb") as fh:
        fh.write(struct.pack("<3i", len(pairs), len(sh_mol), nmol))
        for i, (cls, s1, s2, f, q) in enumerate(pairs):
            fh.write(struct.pack("<5i", cls, s1, s2, pp_s[i] if split else len(f), len(f)))
        fh.write(struct.pack("<%di" % len(sh_mol), *sh_mol))
        fh.write(struct.pack("<%dd" % len(pairs), *[p[4] for p in pairs]))
        for ids in order:
            fh.write(struct.pack("<i", len(ids)))
            if ids:
                fh.write(struct.pack("<%di" % len(ids), *ids))
        fh.write(struct.pack("<2d3i", cut, fp32cut, 1 if split else 0, caps[0], caps[1]))


def _run(harness, tmp_path, molecules, **kw):
    pairs, sh_mol = _pairs(molecules)
    path = str(tmp_path / "in.bin")
    _write(path, pairs, sh_mol, len(molecules), **kw)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([harness, path], capture_output=True, text=True, timeout=600, env=env)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-3000:]
    assert "failures 0" in r.stdout
    return r.stdout


@pytest.mark.parametrize("cut,fp32cut,prim_rel,caps", [
    (1e-12, 0.0, 1e-4, (16, 0)),
    (1e-12, 1e-4, 1e-4, (4 | 4096, 32)),
    (1e-12, 1e-4, 1e-1, (32, 8)),
    (0.0, 0.0, 0.0, (0, 0)),
    (1e-6, 1e-3, 1.0, (1, 1)),
])
def test_worklists_one_cluster(harness, tmp_path, cut, fp32cut, prim_rel, caps):
    from pyquante_amd.molecule import water_cluster
    out = _run(harness, tmp_path, [water_cluster((2, 2, 1), seed=3)], cut=cut, fp32cut=fp32cut,
               prim_rel=prim_rel, split=True, caps=caps)
    assert out.count("class ") == 21


def test_worklists_batch_and_stored_layout(harness, tmp_path):
    """A batch of clusters of uneven size (kets only from the bra's own
    molecule), in the J/K layout and in the stored-integral layout (every
    primitive, no ket split, no runs)."""
    from pyquante_amd.molecule import water_cluster
    mols = [water_cluster(nwater=n, seed=10 + n) for n in (1, 3, 2)]
    _run(harness, tmp_path, mols, cut=1e-12, fp32cut=1e-4, prim_rel=1e-2, split=True, caps=(8, 4))
    _run(harness, tmp_path, mols, cut=1e-12, fp32cut=0.0, prim_rel=0.0, split=False, caps=(0, 0))
