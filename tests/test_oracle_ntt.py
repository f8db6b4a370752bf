"""Pins the oracle's negacyclic NTT (CPU only): brute-force evaluation at the
bit-reversed odd powers of psi, the inverse, and the convolution theorem against
the reference's schoolbook negacyclic product (refPolyMul128,
include/Utils.cuh:236)."""
import random

import numpy as np
import pytest

import oracle_lib as orc
from oracle_lib import lib


def tables(q, log_n):
    n = 1 << log_n
    psi = lib.or_mth_root(2 * n, q)
    t = [np.zeros(n, dtype=np.uint64) for _ in range(4)]
    lib.or_fill_tables(q, psi, log_n, *[orc.P(x) for x in t])
    return psi, t[0], t[1]


@pytest.mark.parametrize("log_n", [4, 9, 12])
def test_forward_is_evaluation_at_bitreversed_odd_powers(log_n):
    q = orc.chain_60bit(2, 1, 1)[0][1]
    n = 1 << log_n
    psi, f, _ = tables(q, log_n)
    rng = np.random.default_rng(log_n)
    a = rng.integers(0, q, size=n, dtype=np.uint64)
    y = a.copy()
    lib.or_ntt_fwd(orc.P(y), log_n, q, orc.P(f))
    ai = [int(v) for v in a]
    for j in random.Random(log_n).sample(range(n), min(n, 24)):
        e = 2 * int(lib.or_bitrev(j, log_n)) + 1
        w = pow(psi, e, q)
        want, p = 0, 1
        for c in ai:
            want = (want + c * p) % q
            p = p * w % q
        assert int(y[j]) == want, j


@pytest.mark.parametrize("log_n", [3, 11, 16])
def test_inverse_undoes_forward(log_n):
    q = orc.chain_60bit(1, 1, 3)[0][0]
    n = 1 << log_n
    _, f, iv = tables(q, log_n)
    a = np.random.default_rng(1).integers(0, q, size=n, dtype=np.uint64)
    y = a.copy()
    lib.or_ntt_fwd(orc.P(y), log_n, q, orc.P(f))
    lib.or_ntt_inv(orc.P(y), log_n, q, orc.P(iv))
    assert np.array_equal(y, a)


@pytest.mark.parametrize("q", [65537, 0x7FFE28001, (1 << 60) - (1 << 18) + 1])
def test_pointwise_product_is_negacyclic_convolution(q):
    log_n = 7
    n = 1 << log_n
    _, f, iv = tables(q, log_n)
    rng = np.random.default_rng(q % 1000)
    a = rng.integers(0, q, size=n, dtype=np.uint64)
    b = rng.integers(0, q, size=n, dtype=np.uint64)
    c = np.zeros(n, dtype=np.uint64)
    lib.or_negacyclic_mul_schoolbook(orc.P(a), orc.P(b), orc.P(c), n, q)
    # independent check of the schoolbook routine itself
    ai, bi = [int(x) for x in a], [int(x) for x in b]
    want = [0] * n
    for i in range(n):
        for j in range(n):
            k = i + j
            if k < n:
                want[k] += ai[i] * bi[j]
            else:
                want[k - n] -= ai[i] * bi[j]
    assert [int(x) for x in c] == [w % q for w in want]
    fa, fb = a.copy(), b.copy()
    lib.or_ntt_fwd(orc.P(fa), log_n, q, orc.P(f))
    lib.or_ntt_fwd(orc.P(fb), log_n, q, orc.P(f))
    prod = np.array([int(x) * int(y) % q for x, y in zip(fa, fb)], dtype=np.uint64)
    lib.or_ntt_inv(orc.P(prod), log_n, q, orc.P(iv))
    assert np.array_equal(prod, c)
