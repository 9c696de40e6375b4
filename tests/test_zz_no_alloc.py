"""pq_digest_stored_noalloc (J/K over a stored integral array with the
library's static scratch, no allocation) against pq_digest_stored, called
through the C ABI twice in a row, the second time on a smaller basis."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _call(fn, ints, D):
    from pyquante_amd._abi import check
    n = D.shape[0]
    J = torch.full_like(D, np.nan)
    K = torch.full_like(D, np.nan)
    check(fn(n, ctypes.c_void_p(ints.data_ptr()), ctypes.c_uint64(ints.numel()), ctypes.c_void_p(D.data_ptr()),
             ctypes.c_void_p(J.data_ptr()), ctypes.c_void_p(K.data_ptr()),
             ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
    torch.cuda.synchronize()
    return J, K


def test_noalloc_matches_allocating_digest():
    from pyquante_amd._abi import PQError, lib
    from pyquante_amd.basis import BasisSet
    from pyquante_amd.engine import Engine
    from pyquante_amd.molecule import Molecule, h2o_sto3g_config
    r = 1 / 0.52918
    big = Molecule("h2o", [(8, (0, 0, 0)), (1, (r, 0, 0)), (1, (0, r, 0))])
    rng = np.random.default_rng(3)
    for mol, basis in ((big, "6-31g**"), (h2o_sto3g_config(), "sto-3g")):
        eng = Engine(BasisSet(mol, basis), mol)
        ints = eng.get2ints()
        n = eng.nbf
        C = rng.standard_normal((n, n))
        D = torch.as_tensor(C + C.T, device=eng.device)
        print("nbf %d: %d stored integrals" % (n, ints.numel()), flush=True)
        Ja, Ka = _call(lib.pq_digest_stored, ints, D)
        Jn, Kn = _call(lib.pq_digest_stored_noalloc, ints, D)
        dj = torch.max(torch.abs(Ja - Jn)).item()
        dk = torch.max(torch.abs(Ka - Kn)).item()
        print("  max |dJ| %.2e, |dK| %.2e" % (dj, dk), flush=True)
        assert dj < 1e-12 and dk < 1e-12
    # beyond the scratch size: a clear error, nothing launched
    with pytest.raises(PQError, match="no-allocation limit"):
        D = torch.zeros((513, 513), dtype=torch.float64, device="cuda")
        npair = 513 * 514 // 2
        ints = torch.zeros(1, dtype=torch.float64, device="cuda")
        from pyquante_amd._abi import check
        check(lib.pq_digest_stored_noalloc(513, ctypes.c_void_p(ints.data_ptr()),
                                           ctypes.c_uint64(npair * (npair + 1) // 2), ctypes.c_void_p(D.data_ptr()),
                                           None, None, None))
