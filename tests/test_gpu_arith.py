"""GPU parity of the ciphertext arithmetic entry points (negate, scalar
multiply with and without accumulate, plaintext addition) against the oracle,
batched and in place."""
import numpy as np
import pytest

import oracle_lib as orc
import params as P

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=["c1_n4096", "ref14_k2"])
def setup(request):
    import kunpbfv_amd as kb
    d = {"c1_n4096": P.c1_params(), "ref14_k2": P.REF14}[request.param]
    ctx = kb.Context(d["log_n"], d["q"], d["p"], d["t"], d["gamma"], d["plain"])
    oc = orc.Ctx(d["log_n"], d["q"], d["p"], d["t"], d["gamma"], d["plain"])
    sk = ctx.keygen_sk(bytes([1]) * 32)
    pk = ctx.keygen_pk(bytes([2]) * 32, sk)
    return kb, ctx, oc, sk, pk


def test_arith_matches_oracle(setup):
    kb, ctx, oc, sk, pk = setup
    rng = np.random.default_rng(4)
    batch = 3
    m1 = rng.integers(0, ctx.plain, size=(batch, ctx.n), dtype=np.uint64)
    m2 = rng.integers(0, ctx.plain, size=(batch, ctx.n), dtype=np.uint64)
    d1, d2 = kb.to_device(m1), kb.to_device(m2)
    ctx.encode(d1)
    ctx.encode(d2)
    c1 = ctx.encrypt(bytes([3]) * 32, pk, d1)
    c2 = ctx.encrypt(bytes([4]) * 32, pk, d2)
    h1, h2, e2 = kb.to_host(c1), kb.to_host(c2), kb.to_host(d2)
    neg = kb.to_host(ctx.negate(c1))
    ms = kb.to_host(ctx.mul_scalar(c1, 987654321))
    msa = kb.to_host(ctx.mul_scalar(c1, 3, add=c2))
    ap = kb.to_host(ctx.add_plain(c1, d2))
    for b in range(batch):
        assert np.array_equal(neg[b], oc.negate(h1[b]))
        assert np.array_equal(ms[b], oc.mul_scalar(h1[b], 987654321))
        assert np.array_equal(msa[b], oc.mul_scalar(h1[b], 3, add=h2[b]))
        assert np.array_equal(ap[b], oc.add_plain(h1[b], e2[b]))
    # in place
    ctx.negate(c1, out=c1)
    assert np.array_equal(kb.to_host(c1), neg)
    dec = ctx.decrypt(sk, ctx.add_plain(c1, d2))
    ctx.decode(dec)
    want = (m2.astype(object) - m1.astype(object)) % ctx.plain
    assert np.array_equal(kb.to_host(dec).astype(object), want)


def test_plain_mult_matches_oracle(setup):
    """Plaintext x ciphertext (BFVScheme::multConstAndEqual(Ciphertext&,
    Plaintext&), include/BFVScheme.cuh:226-245): the plaintext lifted into
    every Q limb and taken to the NTT domain (encode_ntt), then a pointwise
    product with both components.  Equal to the oracle's NTT of the lifted
    plaintext times the ciphertext mod q_i, and the product decrypts to the
    slot-wise product."""
    kb, ctx, oc, sk, pk = setup
    rng = np.random.default_rng(5)
    batch = 2
    m1 = rng.integers(0, ctx.plain, size=(batch, ctx.n), dtype=np.uint64)
    m2 = rng.integers(0, ctx.plain, size=ctx.n, dtype=np.uint64)
    d1, d2 = kb.to_device(m1), kb.to_device(m2[None].copy())
    ctx.encode(d1)
    ctx.encode(d2)
    c1 = ctx.encrypt(bytes([6]) * 32, pk, d1)
    e2 = kb.to_host(d2)[0]
    lifted = np.broadcast_to(e2, (ctx.qn, ctx.n)).copy()
    pt = kb.to_device(lifted)
    ctx.ntt(pt, pt, +1, ctx.oQ, ctx.qn)
    out = ctx.empty(*c1.shape)
    for b in range(batch):
        for h in range(2):
            ctx.poly_op("mul", c1[b, h], pt, out[b, h], ctx.oQ, ctx.qn)
    got = kb.to_host(out)
    o_pt = oc.ntt_fwd(lifted, oc.oQ, oc.qn)
    assert np.array_equal(kb.to_host(pt), o_pt)
    q = [int(v) for v in oc.mod[oc.oQ:oc.oQ + oc.qn]]
    h1 = kb.to_host(c1)
    for b in range(batch):
        for h in range(2):
            for i in range(ctx.qn):
                want = (h1[b, h, i].astype(object) * o_pt[i].astype(object)) % q[i]
                assert np.array_equal(got[b, h, i].astype(object), want), (b, h, i)
    dec = ctx.decrypt(sk, out)
    ctx.decode(dec)
    want = (m1.astype(object) * m2[None].astype(object)) % ctx.plain
    assert np.array_equal(kb.to_host(dec).astype(object), want)
