"""GPU parity of Galois automorphisms, Galois keys and rotation through the C
ABI: bit-exact with the oracle (automorphism, key, rotated ciphertexts for
batches of 3 and of 20, the latter through the matrix-core inner product), and
the rotated ciphertexts decrypt to the permuted slots."""
import numpy as np
import pytest

import oracle_lib as orc
import params as P

pytestmark = pytest.mark.gpu

SETS = {
    "c1_n4096": P.c1_params(),
    "chain60_n8192": P.chain60(13, 3, 1, 3, 1),
    "ref14_k2": P.REF14,
}


@pytest.fixture(scope="module", params=list(SETS))
def setup(request):
    import kunpbfv_amd as kb
    d = SETS[request.param]
    ctx = kb.Context(d["log_n"], d["q"], d["p"], d["t"], d["gamma"], d["plain"])
    oc = orc.Ctx(d["log_n"], d["q"], d["p"], d["t"], d["gamma"], d["plain"])
    sk = ctx.keygen_sk(bytes([1]) * 32)
    pk = ctx.keygen_pk(bytes([2]) * 32, sk)
    return kb, ctx, oc, sk, pk


def test_galois_elt_and_automorphism(setup):
    kb, ctx, oc, _, _ = setup
    for steps in (0, 1, 7, -1):
        assert ctx.galois_elt(steps) == oc.galois_elt(steps)
    rng = np.random.default_rng(3)
    pq = ctx.pn + ctx.qn
    x = np.stack([rng.integers(0, int(q), size=(2, ctx.n), dtype=np.uint64) for q in ctx.mod[:pq]], axis=1)
    x[:, :, :5] = 0  # zeros stay zero under negation
    xd = kb.to_device(np.ascontiguousarray(x))
    out = ctx.empty(*x.shape)
    for g in (5, ctx.galois_elt(3), 2 * ctx.n - 1):
        ctx.automorphism(g, xd, out, ctx.oP, pq, polys=2)
        assert np.array_equal(kb.to_host(out), oc.automorphism(g, x, ctx.oP, pq, 2)), g


@pytest.mark.parametrize("steps", [1, -1])
def test_galois_key_matches_oracle(setup, steps):
    kb, ctx, oc, sk, _ = setup
    g = ctx.galois_elt(steps)
    key = ctx.keygen_galois(bytes([9]) * 32, sk, g)
    assert np.array_equal(kb.to_host(key), oc.keygen_galois(bytes([9]) * 32, kb.to_host(sk), g))


@pytest.mark.parametrize("batch", [3, 20])
def test_rotate_matches_oracle_and_decrypts(setup, batch):
    kb, ctx, oc, sk, pk = setup
    g = ctx.galois_elt(1)
    key = ctx.keygen_galois(bytes([9]) * 32, sk, g)
    rng = np.random.default_rng(batch)
    m = rng.integers(0, ctx.plain, size=(batch, ctx.n), dtype=np.uint64)
    md = kb.to_device(m)
    ctx.encode(md)
    ct = ctx.encrypt(bytes([10]) * 32, pk, md)
    rot = ctx.rotate(key, g, ct)
    cth, roth = kb.to_host(ct), kb.to_host(rot)
    okey = kb.to_host(key)
    for b in (0, batch - 1):
        assert np.array_equal(roth[b], oc.rotate(okey, g, cth[b])), b
    dec = ctx.decrypt(sk, rot)
    ctx.decode(dec)
    got = kb.to_host(dec)
    if oc.batching:
        n, m2 = ctx.n, 2 * ctx.n
        idx = {(2 * int(orc.lib.or_bitrev(j, oc.log_n)) + 1) % m2: j for j in range(n)}
        perm = np.array([idx[(2 * int(orc.lib.or_bitrev(j, oc.log_n)) + 1) * g % m2] for j in range(n)])
        assert np.array_equal(got, m[:, perm])
    else:
        for b in range(batch):
            assert np.array_equal(got[b], oc.automorphism(g, m[b][None], oc.oPlain, 1)[0])
    # in place (out aliases ct) gives the same bits
    ctx.rotate(key, g, ct, out=ct)
    assert np.array_equal(kb.to_host(ct), roth)


def test_seal_order_rotation_is_cyclic(setup):
    kb, ctx, oc, sk, pk = setup
    if not oc.batching:
        pytest.skip("SEAL slot order needs batching")
    rng = np.random.default_rng(21)
    batch, row = 2, ctx.n // 2
    v = rng.integers(0, ctx.plain, size=(batch, ctx.n), dtype=np.uint64)
    vd = kb.to_device(v)
    ctx.encode_seal(vd)
    enc = kb.to_host(vd)
    for b in range(batch):
        assert np.array_equal(enc[b], oc.encode_seal(v[b]))
    g = ctx.galois_elt(1)
    key = ctx.keygen_galois(bytes([12]) * 32, sk, g)
    ct = ctx.encrypt(bytes([13]) * 32, pk, vd)
    dec = ctx.decrypt(sk, ctx.rotate(key, g, ct))
    ctx.decode_seal(dec)
    got = kb.to_host(dec)
    want = np.concatenate([np.roll(v[:, :row], -1, axis=1), np.roll(v[:, row:], -1, axis=1)], axis=1)
    assert np.array_equal(got, want)
