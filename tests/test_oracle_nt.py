"""Pins the oracle's scalar number theory and twiddle tables (CPU only).

Checks against Python big-integer arithmetic, against the values the
reference's own parameter sets fix (tests/golden/ref_chains.json), and against
the defining properties of the roots."""
import json
import os
import random

import numpy as np

import oracle_lib as orc
from oracle_lib import lib

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def golden(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def test_mulmod_powmod_inv_shoup_against_python():
    rng = random.Random(1)
    qs = [(1 << 60) - (1 << 18) + 1, 65537, 0x7FFE28001, 36028797001138177]
    for _ in range(2000):
        q = rng.choice(qs)
        a, b = rng.randrange(q), rng.randrange(q)
        assert lib.or_mulmod(a, b, q) == a * b % q
        e = rng.randrange(1 << 62)
        assert lib.or_powmod(a, e, q) == pow(a, e, q)
        if a:
            assert lib.or_invmod(a, q) == pow(a, -1, q)
        assert lib.or_shoup(a, q) == (a << 64) // q
    for a, b, q in [(0, 0, 3), (2, 2, 3), (qs[0] - 1, qs[0] - 1, qs[0])]:
        assert lib.or_mulmod(a, b, q) == a * b % q


def test_barrett_single_any_128bit_value():
    rng = random.Random(2)
    for _ in range(2000):
        q = rng.randrange(3, 1 << 60) | 1
        v = rng.randrange(1 << 128)
        assert lib.or_barrett_single(v & ((1 << 64) - 1), v >> 64, q) == v % q


def _is_prime_py(n):
    if n < 2:
        return False
    small = [2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37]
    for p in small:
        if n % p == 0:
            return n == p
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for a in small:
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(s - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def test_is_prime_matches_trial_division_and_mr():
    sieve = np.ones(20000, dtype=bool)
    sieve[:2] = False
    for i in range(2, 150):
        sieve[i * i::i] = False
    for n in range(20000):
        assert bool(lib.or_is_prime(n)) == bool(sieve[n]), n
    rng = random.Random(3)
    for _ in range(300):
        n = rng.randrange(1 << 59, 1 << 60)
        assert bool(lib.or_is_prime(n)) == _is_prime_py(n)


def test_reference_t_chain_is_the_60bit_rule():
    # include/Params.cuh:129-140 lists the reference's T primes
    g = golden("ref_chains.json")
    assert orc.primes_descending(1 << 60, 1 << 17, 4) == g["ref_t"]


def test_reference_r_basis_of_logn15_chain():
    # R basis the reference derives with get_primes_below (include/BFVcontext.cuh:57)
    g = golden("ref_chains.json")
    qmin = min(g["ref15_q"])
    out = np.zeros(len(g["ref15_q"]) + 1, dtype=np.uint64)
    assert lib.or_primes_below(1 << 15, qmin, len(out), orc.P(out)) == 0
    assert int(out[0]) == g["ref15_r_first"] and int(out[-1]) == g["ref15_r_last"]
    for r in out:
        assert _is_prime_py(int(r)) and (int(r) - 1) % (1 << 16) == 0


def test_primitive_roots_and_psi():
    g = golden("ref_chains.json")
    assert lib.or_primitive_root(g["ref15_q"][0]) == g["ref15_q0_generator"]
    assert lib.or_mth_root(1 << 16, g["ref15_q"][0]) == g["ref15_q0_psi_n32768"]
    assert lib.or_primitive_root(g["ref_t"][0]) == g["ref_t0_generator"]
    # defining properties on a range of chain primes
    q, p, t = orc.chain_60bit(6, 1, 4)
    for m in q + p + t:
        gen = lib.or_primitive_root(m)
        # gen has full order: gen^((m-1)/f) != 1 for the prime factors f of m-1
        x, fac, f = m - 1, set(), 2
        while f * f <= x and f < 1 << 22:
            while x % f == 0:
                fac.add(f)
                x //= f
            f += 1
        if x > 1:
            assert _is_prime_py(x)
            fac.add(x)
        assert all(pow(gen, (m - 1) // f, m) != 1 for f in fac)
        # and it is the smallest such element
        for r in range(2, gen):
            assert any(pow(r, (m - 1) // f, m) == 1 for f in fac)
        for log_n in (12, 16):
            psi = lib.or_mth_root(2 << log_n, m)
            assert pow(psi, 1 << log_n, m) == m - 1


def test_twiddle_tables_bitreversed_with_folded_ninv():
    q = orc.chain_60bit(1, 1, 1)[0][0]
    log_n = 10
    n = 1 << log_n
    psi = lib.or_mth_root(2 * n, q)
    tabs = [np.zeros(n, dtype=np.uint64) for _ in range(4)]
    lib.or_fill_tables(q, psi, log_n, *[orc.P(t) for t in tabs])
    f, iv, fs, ivs = tabs
    psiinv = pow(psi, -1, q)
    for i in range(n):
        j = int(lib.or_bitrev(i, log_n))
        assert int(f[j]) == pow(psi, i, q)
        want = pow(psiinv, i, q) * (pow(n, -1, q) if j == 1 else 1) % q
        assert int(iv[j]) == want
        assert int(fs[j]) == (int(f[j]) << 64) // q
        assert int(ivs[j]) == (int(iv[j]) << 64) // q
