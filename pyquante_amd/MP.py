"""Moller-Plesset MP2 on the GPU (PyQuante/MP.py:20-125).

The reference unpacks the flat unique-ERI array element by element and runs
the four quarter transforms as Python loops of dot products.  Here the array
is unpacked on the device in blocks of rows (pq_unpack_eri) and every quarter
transform is one dense FP64 matrix product on the device:
    (i nu|lam sig) = sum_mu C[mu,i] (mu nu|lam sig)       occupied i, per row block
    (i nu|j sig)   = sum_lam C[lam,j] (i nu|lam sig)
    (i nu|j b)     = sum_sig (i nu|j sig) C[sig,b]
    (i a|j b)      = sum_nu C[nu,a] (i nu|j b)
so the work is O(o n^4) in GEMMs.  Memory: the o n^3 first-quarter tensor
plus one row block of the AO tensor.
"""
import ctypes

import numpy as np
import torch

from . import _abi
from ._abi import check, lib


def _flat_device(Ints):
    from .Ints import Integrals
    if isinstance(Ints, Integrals):
        return Ints.device_array(), Ints.nbf
    if torch.is_tensor(Ints):
        flat = Ints
    else:
        _abi.require_gpu()
        flat = torch.as_tensor(np.ascontiguousarray(np.asarray(Ints, dtype=np.float64)),
                               device=torch.device("cuda", torch.cuda.current_device()))
    npair = int((np.sqrt(8 * flat.numel() + 1) - 1) // 2)
    n = int((np.sqrt(8 * npair + 1) - 1) // 2)
    if npair * (npair + 1) // 2 != flat.numel() or n * (n + 1) // 2 != npair:
        raise ValueError("not a flat unique-ERI array")
    return flat, n


def _unpack_rows(flat, n, mu0, nmu, out):
    check(lib.pq_unpack_eri(n, ctypes.c_void_p(flat.data_ptr()), ctypes.c_uint64(flat.numel()), int(mu0), int(nmu),
                            ctypes.c_void_p(out.data_ptr()),
                            ctypes.c_void_p(torch.cuda.current_stream(flat.device).cuda_stream)))


def half_transform(Ints, Cocc, block_bytes=2 << 30):
    """(i nu|j sig) for the occupied columns Cocc (n x o): an (o, n, o, n)
    device tensor, from the first two quarter transforms."""
    flat, n = _flat_device(Ints)
    dev = flat.device
    Co = torch.as_tensor(Cocc, dtype=torch.float64, device=dev).contiguous()
    o = Co.shape[1]
    n3 = n * n * n
    rows = max(1, min(n, block_bytes // (8 * n3)))
    T1 = torch.zeros((o, n3), dtype=torch.float64, device=dev)
    block = torch.empty((rows, n3), dtype=torch.float64, device=dev)
    for mu0 in range(0, n, rows):
        m = min(rows, n - mu0)
        _unpack_rows(flat, n, mu0, m, block)
        T1.addmm_(Co[mu0:mu0 + m].T, block[:m])
    del block
    # (i nu|j sig) = sum_lam C[lam,j] (i nu|lam sig): batched over (i, nu)
    T2 = torch.matmul(Co.T, T1.view(o * n, n, n))
    return T2.view(o, n, o, n)


def mo_ovov(Ints, orbs, nclosed, nvirt=None, C2=None):
    """(a r|b s) for occupied a, b and the columns r, s of C2 (default: the
    nvirt virtual orbitals): an (o, v, o, v) device tensor."""
    orbs = torch.as_tensor(orbs, dtype=torch.float64)
    n = orbs.shape[0]
    if C2 is None:
        nvirt = orbs.shape[1] - nclosed if nvirt is None else nvirt
        C2 = orbs[:, nclosed:nclosed + nvirt]
    half = half_transform(Ints, orbs[:, :nclosed])
    dev = half.device
    Cv = torch.as_tensor(C2, dtype=torch.float64, device=dev).contiguous()
    o, v = nclosed, Cv.shape[1]
    T3 = torch.matmul(half.view(o * n, o, n), Cv)          # (i nu|j b)
    X = torch.matmul(Cv.T, T3.view(o, n, o * v))            # (i a|j b)
    return X.view(o, v, o, v)


def MP2(aoints, orbs, orbe, nclosed, nvirt):
    """Closed-shell MP2 correlation energy (MP.py:76-99):
    sum_{abrs} (ar|bs)[2(ar|bs) - (as|br)] / (e_a + e_b - e_r - e_s)."""
    X = mo_ovov(aoints, orbs, nclosed, nvirt)
    e = torch.as_tensor(np.asarray(orbe, dtype=np.float64), device=X.device)
    eo, ev = e[:nclosed], e[nclosed:nclosed + nvirt]
    den = eo[:, None, None, None] - ev[None, :, None, None] + eo[None, None, :, None] - ev[None, None, None, :]
    return float(torch.sum(X * (2.0 * X - X.permute(0, 3, 2, 1)) / den))


# the reference's EN2 evaluates the same expression (MP.py:101-125)
EN2 = MP2


def TransformIntsMP2(Ints, orbs, nclosed):
    """MP.py:20-74: the MO integrals (a i|b j), a, b occupied, i, j any MO, in
    the reference's flat layout of length nmo(nmo+1)(nmo^2+nmo+2)/8 (other
    entries zero), as a numpy array."""
    orbs = torch.as_tensor(orbs, dtype=torch.float64)
    nmo = orbs.shape[1]
    X = mo_ovov(Ints, orbs, nclosed, C2=orbs)                  # (a, i, b, j)
    dev = X.device
    a = torch.arange(nclosed, device=dev).view(-1, 1, 1, 1)
    i = torch.arange(nmo, device=dev).view(1, -1, 1, 1)
    b = torch.arange(nclosed, device=dev).view(1, 1, -1, 1)
    j = torch.arange(nmo, device=dev).view(1, 1, 1, -1)

    def pair(p, q):
        hi, lo = torch.maximum(p, q), torch.minimum(p, q)
        return hi * (hi + 1) // 2 + lo
    ai, bj = pair(a, i), pair(b, j)
    idx = pair(ai, bj).expand(X.shape).reshape(-1)
    totlen = nmo * (nmo + 1) * (nmo * nmo + nmo + 2) // 8
    out = torch.zeros(totlen, dtype=torch.float64, device=dev)
    out[idx] = X.reshape(-1)
    return out.cpu().numpy()
