"""Pins the oracle's key-switching conversions in the auxiliary basis T
(CPU only) with exact CRT arithmetic, on a K = 1 chain and on the reference's
logN = 14 set (K = 2 special primes, gamma = 2)."""
import numpy as np
import pytest

import oracle_lib as orc
import params as Pm


def crt(residues, mods):
    M = 1
    for m in mods:
        M *= m
    x = 0
    for r, m in zip(residues, mods):
        Mi = M // m
        x += int(r) * Mi * pow(Mi, -1, m)
    return x % M, M


@pytest.fixture(scope="module", params=["k1", "ref14"])
def oc(request):
    if request.param == "k1":
        q, p, t = orc.chain_60bit(4, 1, 3)
        return orc.Ctx(12, q, p, t, 1, 65537)
    d = Pm.REF14
    return orc.Ctx(12, d["q"], d["p"], d["t"], d["gamma"], d["plain"])


def mods(oc, lo, n):
    return [int(x) for x in oc.mod[lo:lo + n]]


def test_modup_qj_to_t(oc):
    rng = np.random.default_rng(1)
    Q, T = mods(oc, oc.oQ, oc.qn), mods(oc, oc.oT, oc.tn)
    x = np.stack([rng.integers(0, q, size=oc.n, dtype=np.uint64) for q in Q])
    out = oc.modup_qj_to_t(x)
    for d in range(oc.dnum):
        blk = Q[d * oc.pn:(d + 1) * oc.pn]
        for k in range(0, oc.n, 31):
            X, Qj = crt(x[d * oc.pn:d * oc.pn + len(blk), k], blk)
            got = [int(v) for v in out[d, :, k]]
            if oc.pn == 1:
                assert got == [X % t for t in T]  # exact at K = 1
            else:  # float estimate: X or X + Qj, both congruent mod Qj
                assert got in ([X % t for t in T], [(X + Qj) % t for t in T], [(X - Qj) % t for t in T])


def test_modup_pq_to_t(oc):
    rng = np.random.default_rng(2)
    pq = oc.pn + oc.qn
    PQ, T = mods(oc, 0, pq), mods(oc, oc.oT, oc.tn)
    x = np.stack([rng.integers(0, m, size=oc.n, dtype=np.uint64) for m in PQ])
    out = oc.modup_pq_to_t(x)
    for b in range(oc.ri_blocks):
        blk = PQ[b * oc.gamma:(b + 1) * oc.gamma]
        for k in range(0, oc.n, 37):
            X, Rb = crt(x[b * oc.gamma:b * oc.gamma + len(blk), k], blk)
            got = [int(v) for v in out[b, :, k]]
            assert got in ([X % t for t in T], [(X + Rb) % t for t in T], [(X - Rb) % t for t in T])


def test_inner_product_is_mod_t_dot_product(oc):
    rng = np.random.default_rng(3)
    T = mods(oc, oc.oT, oc.tn)
    cin = np.stack([np.stack([rng.integers(0, t, size=oc.n, dtype=np.uint64) for t in T])
                    for _ in range(oc.dnum)])
    key = np.stack([np.stack([np.stack([np.stack([rng.integers(0, t, size=oc.n, dtype=np.uint64) for t in T])
                                        for _ in range(oc.ri_blocks)]) for _ in range(oc.dnum)])
                    for _ in range(2)])
    out = oc.inner_product(cin, key)
    for h in range(2):
        for b in range(oc.ri_blocks):
            for kk, t in enumerate(T):
                for n in range(0, oc.n, 97):
                    want = sum(int(cin[i, kk, n]) * int(key[h, i, b, kk, n]) for i in range(oc.dnum)) % t
                    assert int(out[h, b, kk, n]) == want


def test_modup_t_to_pq_is_centered_crt(oc):
    """Inputs are small centered values (as the real inner product produces);
    the output is their exact residue in every PQ prime."""
    rng = np.random.default_rng(4)
    pq = oc.pn + oc.qn
    PQ, T = mods(oc, 0, pq), mods(oc, oc.oT, oc.tn)
    Tp = 1
    for t in T:
        Tp *= t
    vals = [[[int(rng.integers(-(1 << 62), 1 << 62)) * int(rng.integers(1, 1 << 40)) for _ in range(oc.n)]
             for _ in range(oc.ri_blocks)] for _ in range(2)]
    x = np.array([[[[v % t for v in vals[h][b]] for t in T] for b in range(oc.ri_blocks)] for h in range(2)],
                 dtype=np.uint64)
    out = oc.modup_t_to_pq(x)
    for h in range(2):
        for b in range(oc.ri_blocks):
            for j in range(b * oc.gamma, min((b + 1) * oc.gamma, pq)):
                for k in range(0, oc.n, 23):
                    assert int(out[h, j, k]) == vals[h][b][k] % PQ[j]


def test_moddown(oc):
    rng = np.random.default_rng(5)
    pq = oc.pn + oc.qn
    PQ = mods(oc, 0, pq)
    P, Q = PQ[:oc.pn], PQ[oc.pn:]
    Pp = 1
    for p in P:
        Pp *= p
    x = np.stack([np.stack([rng.integers(0, m, size=oc.n, dtype=np.uint64) for m in PQ]) for _ in range(2)])
    out = oc.moddown(x)
    for h in range(2):
        for k in range(0, oc.n, 19):
            xp, _ = crt(x[h, :oc.pn, k], P)
            for l, q in enumerate(Q):
                xq = int(x[h, oc.pn + l, k])
                got = int(out[h, l, k])
                # (x_q - [x]_P) / P, with [x]_P exact or off by a multiple v*P, 0 <= v < K
                cands = [((xq - xp - v * Pp) * pow(Pp, -1, q)) % q for v in range(oc.pn)]
                assert got in cands
