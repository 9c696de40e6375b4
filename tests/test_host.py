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
