"""GPU parity: every HIP kernel of the HomMult path against the CPU oracle on the
same seeded inputs, through the C ABI (libkunpbfv.so).  Bit-exact equality is
required everywhere (integer work; the double/float estimates are restated
operation for operation)."""
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
def pair(request):
    import kunpbfv_amd as kb
    d = SETS[request.param]
    ctx = kb.Context(d["log_n"], d["q"], d["p"], d["t"], d["gamma"], d["plain"])
    oc = orc.Ctx(d["log_n"], d["q"], d["p"], d["t"], d["gamma"], d["plain"])
    return ctx, oc


def rand_mod(rng, mods, n, polys=1):
    mods = np.asarray(mods, dtype=np.uint64)
    out = np.empty((polys, len(mods), n), dtype=np.uint64)
    for i, q in enumerate(mods):
        out[:, i, :] = rng.integers(0, int(q), size=(polys, n), dtype=np.uint64)
    return out


def test_tables_match_oracle(pair):
    ctx, oc = pair
    assert np.array_equal(ctx.mod, oc.mod)
    names = ["q_hatinv", "q_hat_mod_r", "alphaQ_mod_r", "r_hatinv", "r_hat_mod_q", "alphaR_mod_q",
             "sr_modr", "dec_modt", "dec_B_modt", "tinv_mod_q", "negQ_mod_t", "ri_hatinv",
             "ri_hat_mod_t", "ri_mod_t", "qj_hatinv", "qj_hat_mod_t", "qj_mod_t", "t_hatinv",
             "t_hat_mod_pq", "t_mod_pq", "halfT_mod", "p_hat_mod_q", "p_hatinv", "P_mod_q",
             "Pinv_mod_q", "dec_msb"]
    for name in names:
        a, b = ctx.export(name), oc.export(name)
        # the oracle pads per-block tables of a short last block with zeros
        n = min(len(a), len(b))
        assert np.array_equal(a[:n], b[:n]), name
    for name in ["q_inv", "r_inv", "sr_frac", "dec_frac", "dec_B_frac"]:
        assert np.array_equal(ctx.export_f64(name), oc.export_f64(name)), name
    psi = ctx.export("psi")
    assert np.array_equal(psi[:-1], oc.psi[:-1])


@pytest.mark.parametrize("which", ["Q", "R", "T", "P"])
def test_ntt_roundtrip_matches_oracle(pair, which):
    import kunpbfv_amd as kb
    ctx, oc = pair
    mod0, limbs = {"Q": (ctx.oQ, ctx.qn), "R": (ctx.oR, ctx.rn), "T": (ctx.oT, ctx.tn),
                   "P": (ctx.oP, ctx.pn)}[which]
    rng = np.random.default_rng(11)
    x = rand_mod(rng, ctx.mod[mod0:mod0 + limbs], ctx.n, polys=3)
    xd = kb.to_device(x)
    yd = ctx.empty(*x.shape)
    ctx.ntt(xd, yd, +1, mod0, limbs, polys=3)
    y = kb.to_host(yd)
    assert np.array_equal(y, oc.ntt_fwd(x, mod0, limbs, 3))
    ctx.ntt(yd, yd, -1, mod0, limbs, polys=3)
    assert np.array_equal(kb.to_host(yd), x)


def test_ntt_lazy_input_and_strides(pair):
    """out-of-place NTT with a broadcast source (the fused digit->T path)."""
    import kunpbfv_amd as kb
    ctx, oc = pair
    rng = np.random.default_rng(5)
    q0 = int(ctx.mod[ctx.oQ])
    src = rng.integers(0, q0, size=(1, ctx.n), dtype=np.uint64)
    dst = ctx.empty(ctx.tn, ctx.n)
    N = ctx.n
    ctx.ntt(kb.to_device(src), dst, +1, ctx.oT, ctx.tn, polys=1, batch=1, s=(0, 0, 0), d=(0, 0, N))
    want = oc.ntt_fwd(np.stack([src[0] % ctx.mod[ctx.oT + k] for k in range(ctx.tn)]), ctx.oT, ctx.tn)
    assert np.array_equal(kb.to_host(dst), want)


def test_bconv_matches_oracle(pair):
    import kunpbfv_amd as kb
    ctx, oc = pair
    rng = np.random.default_rng(3)
    x = rand_mod(rng, ctx.mod[ctx.oQ:ctx.oQ + ctx.qn], ctx.n)[0]
    y = ctx.empty(ctx.rn, ctx.n)
    ctx.bconv(kb.to_device(x), y, True)
    assert np.array_equal(kb.to_host(y), oc.bconv_q_to_r(x))
    xr = rand_mod(rng, ctx.mod[ctx.oR:ctx.oR + ctx.rn], ctx.n)[0]
    yq = ctx.empty(ctx.qn, ctx.n)
    ctx.bconv(kb.to_device(xr), yq, False)
    assert np.array_equal(kb.to_host(yq), oc.bconv_r_to_q(xr))


def test_tensor_scale_round_decrypt_match_oracle(pair):
    import kunpbfv_amd as kb
    ctx, oc = pair
    rng = np.random.default_rng(9)
    sn = ctx.qn + ctx.rn
    mods = ctx.mod[ctx.oQ:ctx.oQ + sn]
    c1 = rand_mod(rng, mods, ctx.n, 2)
    c2 = rand_mod(rng, mods, ctx.n, 2)
    out = ctx.empty(3, sn, ctx.n)
    ctx.tensor(kb.to_device(c1), kb.to_device(c2), out, ctx.oQ, sn)
    assert np.array_equal(kb.to_host(out), oc.tensor(c1, c2, sn, ctx.oQ))
    x = rand_mod(rng, mods, ctx.n)[0]
    y = ctx.empty(ctx.rn, ctx.n)
    ctx.scale_round(kb.to_device(x), y)
    assert np.array_equal(kb.to_host(y), oc.scale_round(x))
    xq = rand_mod(rng, ctx.mod[ctx.oQ:ctx.oQ + ctx.qn], ctx.n)[0]
    m = ctx.empty(1, ctx.n)
    ctx.hps_decrypt(kb.to_device(xq), m)
    assert np.array_equal(kb.to_host(m)[0], oc.hps_decrypt(xq))


def test_keyswitch_stages_match_oracle(pair):
    import kunpbfv_amd as kb
    ctx, oc = pair
    rng = np.random.default_rng(13)
    N = ctx.n
    xq = rand_mod(rng, ctx.mod[ctx.oQ:ctx.oQ + ctx.qn], N)[0]
    ct = ctx.empty(ctx.dnum, ctx.tn, N)
    ctx.modup_qj_to_t(kb.to_device(xq), ct)
    assert np.array_equal(kb.to_host(ct), oc.modup_qj_to_t(xq))
    pq = ctx.pn + ctx.qn
    xpq = rand_mod(rng, ctx.mod[:pq], N)[0]
    kt = ctx.empty(ctx.ri_blocks, ctx.tn, N)
    ctx.modup_pq_to_t(kb.to_device(xpq), kt)
    assert np.array_equal(kb.to_host(kt), oc.modup_pq_to_t(xpq))
    tm = ctx.mod[ctx.oT:ctx.oT + ctx.tn]
    cin = np.stack([rand_mod(rng, tm, N)[0] for _ in range(ctx.dnum)])
    key = np.stack([np.stack([np.stack([rand_mod(rng, tm, N)[0] for _ in range(ctx.ri_blocks)])
                              for _ in range(ctx.dnum)]) for _ in range(2)])
    ip = ctx.empty(2, ctx.ri_blocks, ctx.tn, N)
    ctx.inner_product(kb.to_device(cin), kb.to_device(key), ip)
    assert np.array_equal(kb.to_host(ip), oc.inner_product(cin, key))
    xt = np.stack([np.stack([rand_mod(rng, tm, N)[0] for _ in range(ctx.ri_blocks)]) for _ in range(2)])
    pqo = ctx.empty(2, pq, N)
    ctx.modup_t_to_pq(kb.to_device(xt), pqo)
    assert np.array_equal(kb.to_host(pqo), oc.modup_t_to_pq(xt))
    xin = np.stack([rand_mod(rng, ctx.mod[:pq], N)[0] for _ in range(2)])
    md = ctx.empty(2, ctx.qn, N)
    ctx.moddown(kb.to_device(xin), md)
    assert np.array_equal(kb.to_host(md), oc.moddown(xin))


def test_samplers_match_oracle(pair):
    import kunpbfv_amd as kb
    ctx, oc = pair
    seed = bytes(range(100, 132))
    for kind in ["uniform", "ternary", "zo", "cbd"]:
        out = ctx.empty(ctx.qn, ctx.n)
        ctx.sample(kind, seed, 77 << 32, ctx.oQ, ctx.qn, out)
        assert np.array_equal(kb.to_host(out), oc.sample(kind, seed, 77 << 32, ctx.oQ, ctx.qn)), kind


def test_xsalsa20_matches_oracle():
    import kunpbfv_amd as kb
    key, nonce = bytes(range(32)), bytes(range(50, 74))
    for n in (1, 64, 1000, 4096 + 7):
        assert kb.xsalsa20(key, nonce, n) == orc.xsalsa20(key, nonce, n)
    assert kb.xsalsa20(key, nonce, 300, block0=5) == orc.xsalsa20(key, nonce, 300, block0=5)


def _keys(ctx, oc):
    s = [bytes([i]) * 32 for i in (21, 22, 23)]
    sk, pk, rlk = ctx.keygen_sk(s[0]), None, None
    pk = ctx.keygen_pk(s[1], sk)
    rlk = ctx.keygen_rlk(s[2], sk)
    o_sk = oc.keygen_sk(s[0])
    return sk, pk, rlk, o_sk, oc.keygen_pk(s[1], o_sk), oc.keygen_rlk(s[2], o_sk)


def test_scheme_bit_exact_with_oracle(pair):
    import kunpbfv_amd as kb
    ctx, oc = pair
    sk, pk, rlk, o_sk, o_pk, o_rlk = _keys(ctx, oc)
    assert np.array_equal(kb.to_host(sk), o_sk)
    assert np.array_equal(kb.to_host(pk), o_pk)
    assert np.array_equal(kb.to_host(rlk), o_rlk)
    rng = np.random.default_rng(17)
    batch = 3
    m = rng.integers(0, ctx.plain, size=(batch, ctx.n), dtype=np.uint64)
    md = kb.to_device(m)
    ctx.encode(md)
    enc_m = kb.to_host(md)
    for b in range(batch):
        assert np.array_equal(enc_m[b], oc.encode(m[b]))
    seed = bytes([31]) * 32
    ct = ctx.encrypt(seed, pk, md)
    cth = kb.to_host(ct)
    for b in range(batch):
        assert np.array_equal(cth[b], oc.encrypt(seed, o_pk, enc_m[b], entry=b)), b
    dec = ctx.decrypt(sk, ct)
    ctx.decode(dec)
    assert np.array_equal(kb.to_host(dec), m)
    # HomMult: batched (ct[i] * ct[(i+1)%3]) equals the oracle and decrypts right
    c2 = ct[[1, 2, 0]].contiguous()
    prod = ctx.mult_relin(rlk, ct, c2)
    ph = kb.to_host(prod)
    for b in range(batch):
        assert np.array_equal(ph[b], oc.mult_relin(o_rlk, cth[b], cth[(b + 1) % batch])), b
    dec = ctx.decrypt(sk, prod)
    ctx.decode(dec)
    got = kb.to_host(dec)
    for b in range(batch):
        if oc.batching:
            want = (m[b].astype(object) * m[(b + 1) % batch].astype(object)) % ctx.plain
            assert np.array_equal(got[b].astype(object), want), b
    # per-ciphertext launches give the same bits as the batched launch
    one = ctx.mult_relin(rlk, ct[1:2].contiguous(), c2[1:2].contiguous())
    assert np.array_equal(kb.to_host(one)[0], ph[1])


@pytest.mark.parametrize("base", [20, 40])
def test_external_product_matches_oracle(pair, base):
    """base 20 splits the digits inside the NTT's first pass on every set;
    base 40 takes the separate digit kernel where a digit can reach 8 q (the
    36-bit chain of N=4096) and the fused pass elsewhere."""
    import kunpbfv_amd as kb
    ctx, oc = pair
    sk, pk, rlk, o_sk, o_pk, o_rlk = _keys(ctx, oc)
    rng = np.random.default_rng(23)
    mg = np.zeros(ctx.n, dtype=np.uint64)
    mg[3] = 1  # X^3
    rg = ctx.rgsw_encrypt(bytes([41]) * 32, sk, kb.to_device(mg), base)
    o_rg = oc.rgsw_encrypt(bytes([41]) * 32, o_sk, mg, base)
    assert np.array_equal(kb.to_host(rg), o_rg)
    m = rng.integers(0, ctx.plain, size=(2, ctx.n), dtype=np.uint64)
    ct = ctx.encrypt(bytes([42]) * 32, pk, kb.to_device(m))
    out = ctx.external_product(rg, ct, base)
    oh = kb.to_host(out)
    cth = kb.to_host(ct)
    for b in range(2):
        assert np.array_equal(oh[b], oc.external_product(o_rg, cth[b], base)), b
    if base > 20:
        return  # larger digits: the noise bound is not the point here
    dec = kb.to_host(ctx.decrypt(sk, out))
    for b in range(2):
        want = np.zeros(ctx.n, dtype=object)
        mb = m[b].astype(object)
        want[3:] = mb[:-3]
        want[:3] = -mb[-3:]
        assert np.array_equal(dec[b].astype(object) % ctx.plain, want % ctx.plain), b


def test_mult_tensor_and_switching_key_match_oracle(pair):
    """compute_t_QInv_c0c1c2 (kbfv_mult_tensor: the scaled tensor in NTT form)
    and addSWKey_23 (kbfv_keygen_swk) against the oracle; a switching key from
    s^2 relinearises like the relinearisation key."""
    import kunpbfv_amd as kb
    ctx, oc = pair
    s = [bytes([i]) * 32 for i in (51, 52, 53, 54)]
    sk = ctx.keygen_sk(s[0])
    pk = ctx.keygen_pk(s[1], sk)
    rng = np.random.default_rng(55)
    m = rng.integers(0, ctx.plain, size=(3, ctx.n), dtype=np.uint64)
    md = kb.to_device(m)
    ctx.encode(md)
    ct = ctx.encrypt(s[2], pk, md)
    cth = kb.to_host(ct)
    c2 = ct[[1, 2, 0]].contiguous()
    d3 = kb.to_host(ctx.mult_tensor(ct, c2))
    for b, b2 in ((0, 1), (1, 2), (2, 0)):
        want = oc.ntt_fwd(oc.mult_tensor(cth[b], cth[b2]), oc.oQ, oc.qn, 3)
        assert np.array_equal(d3[b], want), b
    # switching key from s^2 (NTT over Q), bit-exact, and usable as a relinearisation key
    s_q = sk[ctx.pn:].contiguous()
    s2 = ctx.empty(ctx.qn, ctx.n)
    ctx.poly_op("mul", s_q, s_q, s2, ctx.oQ, ctx.qn)
    swk = ctx.keygen_swk(s[3], sk, s2)
    o_sk = oc.keygen_sk(s[0])
    o_s2 = kb.to_host(s2)
    assert np.array_equal(kb.to_host(swk), oc.keygen_swk(s[3], o_sk, o_s2))
    prod = ctx.mult_relin(swk, ct[0:1].contiguous(), c2[0:1].contiguous())
    dec = ctx.decrypt(sk, prod)
    ctx.decode(dec)
    got = kb.to_host(dec)[0]
    assert ctx.batching
    assert np.array_equal(got.astype(object), (m[0].astype(object) * m[1].astype(object)) % ctx.plain)
    assert np.array_equal(kb.to_host(prod)[0], oc.mult_relin(kb.to_host(swk), cth[0], cth[1]))
