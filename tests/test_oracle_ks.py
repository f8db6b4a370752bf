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


def exact_block_lift(res, blk):
    """CRT value X in [0, M) of the residues `res` of one block of moduli, and
    the distance of the exact rational sum_j [x_j * (M/m_j)^-1]_{m_j} / m_j to
    the nearest integer (the quantity the float estimate truncates)."""
    from fractions import Fraction
    M = 1
    for m in blk:
        M *= m
    s = Fraction(0)
    for r, m in zip(res, blk):
        rb = int(r) * pow(M // m, -1, m) % m
        s += Fraction(rb, m)
    f = s - (s.numerator // s.denominator)
    X, _ = crt(res, blk)
    return X, min(f, 1 - f)


def check_modup_to_t(oc, which, seed):
    """Q_j -> T (which = "qj") or PQ -> T ("pq"): every output equals the block's
    CRT value in [0, M_b) mod each t_k, exactly, unless the exact overflow sum
    lies within 2^-18 of an integer (float estimate error < 2^-21 at <= 4 terms).
    A rounded instead of truncated estimate gives X - M_b on about half of the
    coefficients (include/BasisConv.cuh:32,112)."""
    rng = np.random.default_rng(seed)
    T = mods(oc, oc.oT, oc.tn)
    if which == "qj":
        src, width, blocks = mods(oc, oc.oQ, oc.qn), oc.pn, oc.dnum
        x = np.stack([rng.integers(0, q, size=oc.n, dtype=np.uint64) for q in src])
        out = oc.modup_qj_to_t(x)
    else:
        src, width, blocks = mods(oc, 0, oc.pn + oc.qn), oc.gamma, oc.ri_blocks
        x = np.stack([rng.integers(0, m, size=oc.n, dtype=np.uint64) for m in src])
        out = oc.modup_pq_to_t(x)
    checked = skipped = 0
    for b in range(blocks):
        blk = src[b * width:(b + 1) * width]
        for k in range(0, oc.n, 7):
            X, dist = exact_block_lift(x[b * width:b * width + len(blk), k], blk)
            if len(blk) > 1 and dist < 2.0 ** -18:
                skipped += 1
                continue
            assert [int(v) for v in out[b, :, k]] == [X % t for t in T], (which, b, k)
            checked += 1
    assert checked > 200 and skipped * 100 < checked
    return checked


def test_modup_qj_to_t(oc):
    check_modup_to_t(oc, "qj", 1)


def test_modup_pq_to_t(oc):
    check_modup_to_t(oc, "pq", 2)


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


def check_moddown(oc, seed=5):
    """out = (x_q - S) * P^-1 mod q exactly, S = x_p mod q at K = 1 and, at
    K > 1, S = sum_i [x_i * (P/p_i)^-1]_{p_i} * (P/p_i): the reference's
    multi-prime kernel accumulates that sum in 128 bits and reduces it mod q
    with no overflow correction (include/BasisConv.cuh:258-298), so S may
    exceed [x]_P by up to (K-1) P and the result is still one exact value."""
    rng = np.random.default_rng(seed)
    pq = oc.pn + oc.qn
    PQ = mods(oc, 0, pq)
    P, Q = PQ[:oc.pn], PQ[oc.pn:]
    Pp = 1
    for p in P:
        Pp *= p
    x = np.stack([np.stack([rng.integers(0, m, size=oc.n, dtype=np.uint64) for m in PQ]) for _ in range(2)])
    out = oc.moddown(x)
    over = 0
    for h in range(2):
        for k in range(0, oc.n, 7):
            S = sum((int(x[h, i, k]) * pow(Pp // p, -1, p) % p) * (Pp // p) for i, p in enumerate(P))
            over += S >= Pp
            for l, q in enumerate(Q):
                want = ((int(x[h, oc.pn + l, k]) - S) * pow(Pp, -1, q)) % q
                assert int(out[h, l, k]) == want, (h, l, k)
    if oc.pn > 1:
        assert over > 0  # the sum exceeded P somewhere: an added correction would show


def test_moddown(oc):
    check_moddown(oc)
