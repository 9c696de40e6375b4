"""The fixed-point accumulation of the deterministic J/K build
(pyquante_amd/csrc/pq_fixed.h), on the CPU through a host build of the same
header: exact splitting, order independence and accuracy against an exactly
rounded sum (math.fsum)."""
import ctypes
import math
import os
import subprocess
from fractions import Fraction

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def fx(tmp_path_factory):
    so = str(tmp_path_factory.mktemp("fx") / "libfx.so")
    subprocess.run(["g++", "-O2", "-shared", "-fPIC", "-o", so, os.path.join(HERE, "native", "fixed_harness.cpp")],
                   check=True)
    lib = ctypes.CDLL(so)
    lib.fx_split3.argtypes = [ctypes.c_double, ctypes.POINTER(ctypes.c_longlong)]
    lib.fx_sum.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
    lib.fx_sum.restype = ctypes.c_double
    lib.fx_value3.argtypes = [ctypes.c_longlong] * 3
    lib.fx_value3.restype = ctypes.c_double
    return lib


def split(fx, x):
    out = (ctypes.c_longlong * 3)()
    fx.fx_split3(x, out)
    return list(out)


def fsum_fx(fx, x, order):
    x = np.ascontiguousarray(x, dtype=np.float64)
    order = np.ascontiguousarray(order, dtype=np.int32)
    return fx.fx_sum(x.ctypes.data, order.ctypes.data, len(x))


def wide_values(rng, n):
    """J/K-like contributions: both signs, magnitudes from 1e-22 to 1e3."""
    return rng.standard_normal(n) * 10.0 ** rng.uniform(-22, 3, n)


def test_split_is_exact_to_2_pow_minus_65(fx):
    rng = np.random.default_rng(5)
    xs = list(wide_values(rng, 3000)) + [0.0, -0.0, 1.0, -1.0, 2.0 ** -64, -2.0 ** -65, 3.5, -1e-25, 2.0 ** 40 + 0.5,
                                         -(2.0 ** 31) - 2.0 ** -30, 0.1, -0.1]
    for x in xs:
        l0, l1, l2 = split(fx, x)
        assert abs(l1) < 2 ** 32 and abs(l0) <= 2 ** 32
        # every limb carries the sign of x
        assert all(v == 0 or (v > 0) == (x > 0) for v in (l0, l1, l2))
        rep = Fraction(l2) + Fraction(l1, 2 ** 32) + Fraction(l0, 2 ** 64)
        assert abs(rep - Fraction(x)) <= Fraction(1, 2 ** 65), x


def test_sum_independent_of_order_and_accurate(fx):
    rng = np.random.default_rng(7)
    x = wide_values(rng, 200000)
    ref = math.fsum(x)
    sums = [fsum_fx(fx, x, rng.permutation(len(x))) for _ in range(4)]
    sums.append(fsum_fx(fx, x, np.arange(len(x))[::-1]))
    # bit-identical whatever the order
    assert len(set(s.hex() for s in sums)) == 1
    # each contribution rounded to 2^-64, then two roundings in the conversion
    assert abs(sums[0] - ref) <= len(x) * 2.0 ** -65 + 2 * math.ulp(ref)


def test_cancellation_and_small_results_keep_relative_precision(fx):
    rng = np.random.default_rng(11)
    big = rng.standard_normal(5000) * 100.0
    small = rng.standard_normal(50) * 1e-9
    x = np.concatenate([big, -big, small])
    got = fsum_fx(fx, x, rng.permutation(len(x)))
    ref = math.fsum(x)          # = sum(small), exactly representable cancellation
    assert abs(got - ref) <= len(x) * 2.0 ** -65 + 2 * math.ulp(ref)
    # exact cancellation gives +0
    y = np.concatenate([big, -big])
    assert fsum_fx(fx, y, rng.permutation(len(y))) == 0.0
    # a negative total below one unit of the integer limb
    z = np.array([-3e-15, 1e-15, -2.5e-17])
    assert abs(fsum_fx(fx, z, np.arange(3)) - math.fsum(z)) <= 3 * 2.0 ** -65 + 2 * math.ulp(math.fsum(z))


def test_limb_sums_carry_across_words(fx):
    """Unnormalised limb sums (far beyond 2^32 in the low words, negative
    high word) convert to the value they denote."""
    cases = [(2 ** 40 + 7, -(2 ** 35) + 3, 5), (-(2 ** 50), 2 ** 45, -3), (2 ** 62, 2 ** 62, -(2 ** 20)),
             (-1, 0, 0), (1, -1, 1), (0, 0, -(2 ** 52))]
    for s0, s1, s2 in cases:
        exact = Fraction(s2) + Fraction(s1, 2 ** 32) + Fraction(s0, 2 ** 64)
        got = fx.fx_value3(s0, s1, s2)
        assert abs(Fraction(got) - exact) <= 2 * abs(exact) * Fraction(2) ** -53 + Fraction(1, 2 ** 64), (s0, s1, s2)
