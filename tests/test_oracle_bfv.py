"""Pins the oracle's BFV scheme (CPU only) by its invariants: keys decrypt to
small error, encryption round-trips, the product of two ciphertexts decrypts
to the product of the plaintexts (slot-wise under batching, negacyclic under
coefficient encoding), and the RGSW external product decrypts to m_rgsw * m."""
import numpy as np
import pytest

import oracle_lib as orc
import params as Pm


def centered(v, q):
    v = v.astype(object)
    return np.where(v > q // 2, v - q, v)


def negacyclic(a, b, t):
    n = len(a)
    out = [0] * n
    a = [int(x) for x in a]
    b = [int(x) for x in b]
    for i in range(n):
        if a[i] == 0:
            continue
        for j in range(n):
            k = i + j
            if k < n:
                out[k] += a[i] * b[j]
            else:
                out[k - n] -= a[i] * b[j]
    return np.array([x % t for x in out], dtype=object)


SETS = {
    "batching_n4096": Pm.c1_params(),
    "coeff_enc_t_not_1_mod_2n": dict(log_n=12, **{k: v for k, v in Pm.chain60(12, 3, 1, 3, 1, 257).items()
                                                  if k != "log_n"}),
    "ref14_k2": dict(Pm.REF14, log_n=12),
}


@pytest.fixture(scope="module", params=list(SETS))
def keys(request):
    d = SETS[request.param]
    oc = orc.Ctx(d["log_n"], d["q"], d["p"], d["t"], d["gamma"], d["plain"])
    sk = oc.keygen_sk(bytes([1]) * 32)
    pk = oc.keygen_pk(bytes([2]) * 32, sk)
    rlk = oc.keygen_rlk(bytes([3]) * 32, sk)
    return oc, sk, pk, rlk


def test_secret_is_ternary_and_public_key_small(keys):
    oc, sk, pk, _ = keys
    pq = oc.pn + oc.qn
    s = oc.ntt_inv(sk, 0, pq)
    for i in range(pq):
        v = centered(s[i], int(oc.mod[i]))
        assert set(np.unique(v)) <= {-1, 0, 1}
        assert np.array_equal(v, centered(s[0], int(oc.mod[0])))
    # b + a*s = e with |e| <= 21 (CBD k = 21)
    e = np.zeros((oc.qn, oc.n), dtype=np.uint64)
    for i in range(oc.qn):
        q = int(oc.mod[oc.oQ + i])
        e[i] = [(int(b) + int(a) * int(x)) % q for a, b, x in zip(pk[0, i], pk[1, i], sk[oc.pn + i])]
    e = oc.ntt_inv(e, oc.oQ, oc.qn)
    for i in range(oc.qn):
        assert np.abs(centered(e[i], int(oc.mod[oc.oQ + i]))).max() <= 21


def test_encrypt_decrypt_roundtrip(keys):
    oc, sk, pk, _ = keys
    rng = np.random.default_rng(1)
    for entry in (0, 5):
        m = rng.integers(0, oc.plain, size=oc.n, dtype=np.uint64)
        ct = oc.encrypt(bytes([4]) * 32, pk, oc.encode(m), entry=entry)
        assert np.array_equal(oc.decode(oc.decrypt(sk, ct)), m)


def test_hommult_relin_decrypts_to_product(keys):
    oc, sk, pk, rlk = keys
    rng = np.random.default_rng(2)
    m1 = rng.integers(0, oc.plain, size=oc.n, dtype=np.uint64)
    m2 = rng.integers(0, oc.plain, size=oc.n, dtype=np.uint64)
    if not oc.batching:
        m2[:] = 0
        m2[[0, 3, 77]] = [2, oc.plain - 1, 5]  # sparse: the negacyclic check stays cheap
    c1 = oc.encrypt(bytes([5]) * 32, pk, oc.encode(m1))
    c2 = oc.encrypt(bytes([6]) * 32, pk, oc.encode(m2))
    prod = oc.mult_relin(rlk, c1, c2)
    got = oc.decode(oc.decrypt(sk, prod)).astype(object)
    if oc.batching:
        want = (m1.astype(object) * m2.astype(object)) % oc.plain
    else:
        want = negacyclic(m2, m1, oc.plain)
    assert np.array_equal(got, want)


def test_relin_alone_recovers_s2_times_d2(keys):
    """relin(d2) decrypts (b + a s) to s^2 d2 up to small noise, over Q."""
    oc, sk, _, rlk = keys
    rng = np.random.default_rng(3)
    Q = [int(x) for x in oc.mod[oc.oQ:oc.oQ + oc.qn]]
    d2 = np.stack([rng.integers(0, q, size=oc.n, dtype=np.uint64) for q in Q])
    ks = oc.relin(rlk, d2)
    ks_ntt = oc.ntt_fwd(ks, oc.oQ, oc.qn, 2)
    d2_ntt = oc.ntt_fwd(d2, oc.oQ, oc.qn)
    s = sk[oc.pn:]
    for i, q in enumerate(Q):
        lhs = [(int(b) + int(a) * int(x)) % q for a, b, x in zip(ks_ntt[0, i], ks_ntt[1, i], s[i])]
        rhs = [int(d) * int(x) * int(x) % q for d, x in zip(d2_ntt[i], s[i])]
        diff = np.array([(a - b) % q for a, b in zip(lhs, rhs)], dtype=np.uint64)
        noise = oc.ntt_inv(np.tile(diff, (oc.qn, 1)), oc.oQ, oc.qn)[i]
        assert np.abs(centered(noise, q)).max() < 1 << 40


def test_external_product_decrypts_to_product(keys):
    oc, sk, pk, _ = keys
    rng = np.random.default_rng(4)
    base = 20 if oc.mod[oc.oQ] > 1 << 40 else 12
    mg = np.zeros(oc.n, dtype=np.uint64)
    mg[1] = 1
    mg[2] = oc.plain - 1  # X - X^2
    rg = oc.rgsw_encrypt(bytes([7]) * 32, sk, mg, base)
    m = rng.integers(0, oc.plain, size=oc.n, dtype=np.uint64)
    ct = oc.encrypt(bytes([8]) * 32, pk, m)
    out = oc.external_product(rg, ct, base)
    got = oc.decrypt(sk, out).astype(object) % oc.plain
    want = negacyclic(mg, m, oc.plain)
    assert np.array_equal(got, want)


def test_reference_logn15_gamma4_is_rejected():
    """The reference's own logN=15 set (gamma = 4, four T primes) cannot hold the
    key-switch inner product; gamma = 2 can."""
    d = Pm.REF15
    with pytest.raises(ValueError):
        orc.Ctx(15, d["q"], d["p"], d["t"], 4, d["plain"])
    orc.Ctx(15, d["q"], d["p"], d["t"], 2, d["plain"])


def test_ciphertext_arithmetic_decrypts(keys):
    """negate, scalar multiply (with accumulate) and plaintext addition decrypt
    to -m, k*m (+ m2) and m + m2 (reference BFVScheme.cuh:192-326)."""
    oc, sk, pk, _ = keys
    rng = np.random.default_rng(9)
    t = oc.plain
    m1 = rng.integers(0, t, size=oc.n, dtype=np.uint64)
    m2 = rng.integers(0, t, size=oc.n, dtype=np.uint64)
    c1 = oc.encrypt(bytes([11]) * 32, pk, oc.encode(m1))
    c2 = oc.encrypt(bytes([12]) * 32, pk, oc.encode(m2))
    dec = lambda ct: oc.decode(oc.decrypt(sk, ct)).astype(object)  # noqa: E731
    a1, a2 = m1.astype(object), m2.astype(object)
    assert np.array_equal(dec(oc.negate(c1)), (-a1) % t)
    k = 12345
    assert np.array_equal(dec(oc.mul_scalar(c1, k)), (k * a1) % t)
    assert np.array_equal(dec(oc.mul_scalar(c1, 7, add=c2)), (7 * a1 + a2) % t)
    assert np.array_equal(dec(oc.add_plain(c1, oc.encode(m2))), (a1 + a2) % t)
