"""numpy restatement of PyQuante's closed-shell MP2 (PyQuante/MP.py:20-99):
full 4-index unpacking of the flat unique-ERI array, the four quarter
transforms as einsums, and the pair-energy sum.  TEST INFRASTRUCTURE ONLY;
for small bases (n^4 doubles in memory)."""
import numpy as np



def ao_tensor(ints, n):
    """(mu nu|lam sig) from the flat array (the unpacking of MP.py:38-46),
    with the canonical index of cints.c:413-425 evaluated on index grids."""
    r = np.arange(n, dtype=np.int64)

    def pair(p, q):
        hi, lo = np.maximum(p, q), np.minimum(p, q)
        return hi * (hi + 1) // 2 + lo
    ij = pair(r[:, None], r[None, :])
    idx = pair(ij[:, :, None, None], ij[None, None, :, :])
    return np.asarray(ints)[idx]


def mp2(ints, orbs, orbe, nclosed, nvirt):
    """E(2) = sum_{abrs} (ar|bs)[2(ar|bs) - (as|br)] / (e_a + e_b - e_r - e_s)
    (MP.py:84-99), a, b occupied, r, s virtual."""
    n = orbs.shape[0]
    A = ao_tensor(ints, n)
    co = orbs[:, :nclosed]
    cv = orbs[:, nclosed:nclosed + nvirt]
    X = np.einsum("ma,nr,lb,sq,mnls->arbq", co, cv, co, cv, A, optimize=True)
    eo, ev = orbe[:nclosed], orbe[nclosed:nclosed + nvirt]
    den = eo[:, None, None, None] - ev[None, :, None, None] + eo[None, None, :, None] - ev[None, None, None, :]
    return float(np.sum(X * (2 * X - X.transpose(0, 3, 2, 1)) / den))
