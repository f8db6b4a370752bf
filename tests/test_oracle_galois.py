"""Pins the oracle's Galois automorphisms and rotation (CPU only).

sigma_g: a(X) -> a(X^g) on Z_q[X]/(X^N + 1), g odd.  Pinned by what the
algebra fixes, not by retyping the index map: sigma_g(X) = X^g reduced with
X^N = -1; sigma_g is multiplicative (checked against the schoolbook negacyclic
product); sigma_g o sigma_h = sigma_gh; and in the NTT domain it moves the value
at evaluation point zeta_j = psi^(2 brv(j) + 1) to the point zeta_j^(g^-1),
checked by brute-force evaluation.  Rotation keys and rotation are pinned by
decryption: decrypt(rotate_g(Enc(m))) = sigma_g(m) (coefficient encoding) and,
with batching, the slot permutation the evaluation points give."""
import numpy as np
import pytest

import oracle_lib as orc
import params as Pm


def centered(v, q):
    v = v.astype(object)
    return np.where(v > q // 2, v - q, v)


@pytest.fixture(scope="module")
def small():
    q, p, t = orc.chain_60bit(2, 1, 3)
    return orc.Ctx(12, q, p, t, 1, 65537)


def test_galois_elements(small):
    oc = small
    m2 = 2 * oc.n
    assert oc.galois_elt(0) == 1 and oc.galois_elt(1) == 5 and oc.galois_elt(3) == 125
    assert oc.galois_elt(-1) == m2 - 1
    # 5 generates a cyclic subgroup of order N/2 in (Z/2N)^*
    assert oc.galois_elt(oc.n // 2) == 1 and oc.galois_elt(oc.n // 4) != 1


def test_monomial_and_multiplicativity(small):
    oc = small
    q = int(oc.mod[oc.oQ])
    n = oc.n
    for g in (5, 25, 2 * n - 1, 3 * n + 7):
        x = np.zeros((1, n), dtype=np.uint64)
        x[0, 1] = 1
        sx = oc.automorphism(g, x, oc.oQ, 1)
        want = np.zeros(n, dtype=np.uint64)
        e = g % (2 * n)
        want[e % n] = 1 if e < n else q - 1
        assert np.array_equal(sx[0], want), g
    rng = np.random.default_rng(1)
    a = rng.integers(0, q, size=(1, n), dtype=np.uint64)
    b = np.zeros((1, n), dtype=np.uint64)
    b[0, rng.integers(0, n, size=6)] = rng.integers(0, q, size=6, dtype=np.uint64)

    def mul(x, y):
        out = np.zeros(n, dtype=np.uint64)
        orc.lib.or_negacyclic_mul_schoolbook(orc.P(np.ascontiguousarray(x[0])), orc.P(np.ascontiguousarray(y[0])),
                                             orc.P(out), n, q)
        return out[None]

    for g in (5, 2 * n - 1, 125):
        lhs = oc.automorphism(g, mul(a, b), oc.oQ, 1)
        rhs = mul(oc.automorphism(g, a, oc.oQ, 1), oc.automorphism(g, b, oc.oQ, 1))
        assert np.array_equal(lhs, rhs), g
    # composition
    g, h = 5, 2 * n - 1
    assert np.array_equal(oc.automorphism(h, oc.automorphism(g, a, oc.oQ, 1), oc.oQ, 1),
                          oc.automorphism(g * h % (2 * n), a, oc.oQ, 1))


def test_ntt_domain_point_map(small):
    """NTT(sigma_g a)[j] = a(zeta_j^g): the value at point zeta_j^g of a."""
    oc = small
    n, q = oc.n, int(oc.mod[oc.oQ])
    psi = int(oc.psi[oc.oQ])
    rng = np.random.default_rng(2)
    a = rng.integers(0, q, size=(1, n), dtype=np.uint64)
    g = 25
    sa = oc.ntt_fwd(oc.automorphism(g, a, oc.oQ, 1), oc.oQ, 1)[0]
    coeffs = [int(v) for v in a[0]]
    for j in rng.integers(0, n, size=6):
        z = pow(psi, 2 * int(orc.lib.or_bitrev(int(j), oc.log_n)) + 1, q)
        zg = pow(z, g, q)
        val, p = 0, 1
        for cf in coeffs:
            val = (val + cf * p) % q
            p = p * zg % q
        assert int(sa[j]) == val


SETS = {
    "batching_n4096": Pm.c1_params(),
    "coeff_enc": dict(log_n=12, **{k: v for k, v in Pm.chain60(12, 3, 1, 3, 1, 257).items() if k != "log_n"}),
    "ref14_k2": dict(Pm.REF14, log_n=12),
}


@pytest.fixture(scope="module", params=list(SETS))
def keys(request):
    d = SETS[request.param]
    oc = orc.Ctx(d["log_n"], d["q"], d["p"], d["t"], d["gamma"], d["plain"])
    sk = oc.keygen_sk(bytes([1]) * 32)
    pk = oc.keygen_pk(bytes([2]) * 32, sk)
    return oc, sk, pk


def slot_perm(oc, g):
    """slot j of sigma_g(m) = slot perm[j] of m (evaluation points mod t)."""
    n, t = oc.n, oc.plain
    m2 = 2 * n
    idx = {}
    for j in range(n):
        idx[(2 * int(orc.lib.or_bitrev(j, oc.log_n)) + 1) % m2] = j
    return np.array([idx[(2 * int(orc.lib.or_bitrev(j, oc.log_n)) + 1) * g % m2] for j in range(n)])


@pytest.mark.parametrize("steps", [1, 3, -1])
def test_rotate_decrypts_to_automorphism(keys, steps):
    oc, sk, pk = keys
    g = oc.galois_elt(steps)
    gk = oc.keygen_galois(bytes([5]) * 32, sk, g)
    rng = np.random.default_rng(steps + 10)
    m = rng.integers(0, oc.plain, size=oc.n, dtype=np.uint64)
    ct = oc.encrypt(bytes([6]) * 32, pk, oc.encode(m))
    rot = oc.rotate(gk, g, ct)
    dec = oc.decrypt(sk, rot)
    if oc.batching:
        assert np.array_equal(oc.decode(dec), m[slot_perm(oc, g)])
    else:
        # sigma_g of the plaintext polynomial over Z_t (the automorphism pinned above)
        want = oc.automorphism(g, m[None], oc.oPlain, 1)[0]
        assert np.array_equal(dec % np.uint64(oc.plain), want)


def test_rotation_by_identity_is_keyswitch_only(keys):
    """g = 1: the key switches from s to s; decryption is unchanged."""
    oc, sk, pk = keys
    gk = oc.keygen_galois(bytes([7]) * 32, sk, 1)
    m = np.arange(oc.n, dtype=np.uint64) % oc.plain
    ct = oc.encrypt(bytes([8]) * 32, pk, oc.encode(m))
    assert np.array_equal(oc.decode(oc.decrypt(sk, oc.rotate(gk, 1, ct))), m)


def test_seal_order_makes_rotation_cyclic():
    """In SEAL's slot order the map is a permutation, encode/decode invert each
    other, and sigma_(5^r) rotates each half-row cyclically left by r while
    sigma_(2N-1) swaps the rows (the point of the order)."""
    d = Pm.c1_params()
    oc = orc.Ctx(d["log_n"], d["q"], d["p"], d["t"], d["gamma"], d["plain"])
    n, row = oc.n, oc.n // 2
    mp = oc.seal_index_map()
    assert sorted(int(x) for x in mp) == list(range(n))
    rng = np.random.default_rng(5)
    v = rng.integers(0, oc.plain, size=n, dtype=np.uint64)
    m = oc.encode_seal(v)
    assert np.array_equal(oc.decode_seal(m), v)
    for r in (1, 3):
        g = oc.galois_elt(r)
        rot = oc.decode_seal(oc.automorphism(g, m[None], oc.oPlain, 1)[0])
        want = np.concatenate([np.roll(v[:row], -r), np.roll(v[row:], -r)])
        assert np.array_equal(rot, want), r
    conj = oc.decode_seal(oc.automorphism(2 * n - 1, m[None], oc.oPlain, 1)[0])
    assert np.array_equal(conj, np.concatenate([v[row:], v[:row]]))
