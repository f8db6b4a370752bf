"""GPU parity at the benchmark's own shape: N = 2^16, 28 + 1 60-bit primes,
4 T primes, gamma = 2, t = 65537.  Some code runs only at this shape (the
k_ntt_pass<16, ...> instantiations, the int8-MFMA paths at 2^16 coefficients,
the fused modUp/modDown for tn = 4, gamma = 2, pn = 1); here it is checked bit
for bit against the CPU oracle on seeded inputs: the NTT over every basis, the
keys, encryption and one batched HomMult+relin of two pairs, plus the
decrypted products (closed forms: one factor is a monomial).  The oracle needs
about 20 s for the relinearisation key and 4 s per HomMult at this size."""
import numpy as np
import pytest

import oracle_lib as orc
import params as P

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pair():
    import kunpbfv_amd as kb
    d = P.chain60(16, 28, 1, 4, 2)
    ctx = kb.Context(d["log_n"], d["q"], d["p"], d["t"], d["gamma"], d["plain"])
    oc = orc.Ctx(d["log_n"], d["q"], d["p"], d["t"], d["gamma"], d["plain"])
    assert ctx.n == 65536 and ctx.qn == 28 and ctx.tn == 4 and ctx.gamma == 2
    return ctx, oc


@pytest.mark.parametrize("which,limbs", [("Q", 3), ("R", 2), ("T", 4), ("P", 1)])
def test_ntt_n65536_matches_oracle(pair, which, limbs):
    import kunpbfv_amd as kb
    ctx, oc = pair
    mod0 = {"Q": ctx.oQ + ctx.qn - limbs, "R": ctx.oR, "T": ctx.oT, "P": ctx.oP}[which]
    rng = np.random.default_rng(64)
    x = np.empty((2, limbs, ctx.n), dtype=np.uint64)
    for i in range(limbs):
        x[:, i, :] = rng.integers(0, int(ctx.mod[mod0 + i]), size=(2, ctx.n), dtype=np.uint64)
    xd = kb.to_device(x)
    yd = ctx.empty(*x.shape)
    ctx.ntt(xd, yd, +1, mod0, limbs, polys=2)
    assert np.array_equal(kb.to_host(yd), oc.ntt_fwd(x, mod0, limbs, 2))
    ctx.ntt(yd, yd, -1, mod0, limbs, polys=2)
    assert np.array_equal(kb.to_host(yd), x)


def test_hommult_n65536_matches_oracle(pair):
    import kunpbfv_amd as kb
    ctx, oc = pair
    s = [bytes([i]) * 32 for i in (21, 22, 23)]
    sk = ctx.keygen_sk(s[0])
    pk = ctx.keygen_pk(s[1], sk)
    rlk = ctx.keygen_rlk(s[2], sk)
    o_sk = oc.keygen_sk(s[0])
    o_pk = oc.keygen_pk(s[1], o_sk)
    assert np.array_equal(kb.to_host(sk), o_sk)
    assert np.array_equal(kb.to_host(pk), o_pk)
    o_rlk = oc.keygen_rlk(s[2], o_sk)
    assert np.array_equal(kb.to_host(rlk), o_rlk)

    # t = 65537 is not 1 mod 2N at N = 2^16: no slots, plaintexts are
    # polynomials mod X^N + 1.  m1 = a X^j makes both products closed forms.
    rng = np.random.default_rng(65)
    a, j = 12345, 777
    m = np.zeros((2, ctx.n), dtype=np.uint64)
    m[0] = rng.integers(0, ctx.plain, size=ctx.n, dtype=np.uint64)
    m[1, j] = a
    md = kb.to_device(m)
    ctx.encode(md)
    enc_m = kb.to_host(md)
    seed = bytes([31]) * 32
    ct = ctx.encrypt(seed, pk, md)
    cth = kb.to_host(ct)
    for b in range(2):
        assert np.array_equal(enc_m[b], oc.encode(m[b]))
        assert np.array_equal(cth[b], oc.encrypt(seed, o_pk, enc_m[b], entry=b)), b

    # batched: (ct0 * ct1, ct1 * ct1)
    c2 = ct[[1, 1]].contiguous()
    prod = kb.to_host(ctx.mult_relin(rlk, ct, c2))
    assert np.array_equal(prod[0], oc.mult_relin(o_rlk, cth[0], cth[1]))
    assert np.array_equal(prod[1], oc.mult_relin(o_rlk, cth[1], cth[1]))

    dec = ctx.decrypt(sk, kb.to_device(prod))
    ctx.decode(dec)
    got = kb.to_host(dec).astype(object)
    t = ctx.plain
    # a X^j * m0: negacyclic shift by j, scaled by a
    sh = np.concatenate([(t - m[0, ctx.n - j:].astype(object)) % t, m[0, :ctx.n - j].astype(object)])
    assert np.array_equal(got[0], (a * sh) % t)
    want1 = np.zeros(ctx.n, dtype=object)
    want1[2 * j] = a * a % t
    assert np.array_equal(got[1], want1)
