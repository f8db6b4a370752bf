"""GPU parity for ciphertexts below the full level, through the C ABI:

* kbfv_mod_switch equals the oracle's modulus switch (the reference's
  rescaleToAndEqual) for one and two special primes;
* kbfv_mult_relin_ragged on a batch of uneven RNS lengths (runs of different
  levels packed back to back, one run of 17 so that the int8-MFMA inner
  product reads the full-level key with its full strides) equals the oracle's
  level-l HomMult entry by entry, equals kbfv_mult_relin on each level's view,
  and decrypts to the slot-wise products;
* a level view encrypts with its own public key and decrypts;
* at the benchmark's shape (N = 2^16, 28 + 1 primes) a ragged batch at levels
  27 and 13 equals the oracle on one entry per level.
"""
import numpy as np
import pytest

import oracle_lib as orc
import params as P

pytestmark = pytest.mark.gpu


def _ctxs(d):
    import kunpbfv_amd as kb
    ctx = kb.Context(d["log_n"], d["q"], d["p"], d["t"], d["gamma"], d["plain"])
    oc = orc.Ctx(d["log_n"], d["q"], d["p"], d["t"], d["gamma"], d["plain"])
    return ctx, oc


def _fresh(ctx, n, seed):
    import kunpbfv_amd as kb
    sk = ctx.keygen_sk(bytes([seed]) * 32)
    pk = ctx.keygen_pk(bytes([seed + 1]) * 32, sk)
    rlk = ctx.keygen_rlk(bytes([seed + 2]) * 32, sk)
    rng = np.random.default_rng(seed)
    slots = rng.integers(0, ctx.plain, size=(2 * n, ctx.n), dtype=np.uint64)
    md = kb.to_device(slots)
    ctx.encode(md)
    ct = ctx.encrypt(bytes([seed + 3]) * 32, pk, md)
    return sk, pk, rlk, slots, ct


SETS = {"k1_n8192": P.chain60(13, 5, 1, 4, 2), "k2_ref14": P.REF14}


@pytest.mark.parametrize("name", list(SETS))
def test_mod_switch_matches_oracle(name):
    import kunpbfv_amd as kb
    ctx, oc = _ctxs(SETS[name])
    sk, pk, rlk, slots, ct = _fresh(ctx, 2, 30)
    L = ctx.qn - 1
    for to in range(L - 1, -1, -1):
        got = kb.to_host(ctx.mod_switch(ct, L, to))
        h = kb.to_host(ct)
        for b in range(ct.shape[0]):
            assert np.array_equal(got[b], oc.mod_switch(h[b], L, to)), (name, to, b)
    # two steps equal one
    mid = ctx.mod_switch(ct, L, L - 1)
    assert np.array_equal(kb.to_host(ctx.mod_switch(mid, L - 1, 0)), kb.to_host(ctx.mod_switch(ct, L, 0)))


@pytest.mark.parametrize("name", list(SETS))
def test_ragged_levels_match_oracle(name):
    import torch
    import kunpbfv_amd as kb
    d = SETS[name]
    ctx, oc = _ctxs(d)
    L = ctx.qn - 1
    levels = [L, L, 1, 2, 2, 2, L - 1] if L >= 3 else [L, 1, 1, L]
    if name == "k1_n8192":
        levels += [2] * 17  # one run fills the matrix-core inner product
    n = len(levels)
    sk, pk, rlk, slots, ct = _fresh(ctx, n, 40)
    c1s, c2s = [], []
    for b, lv in enumerate(levels):
        c1s.append(ctx.mod_switch(ct[b:b + 1].contiguous(), L, lv) if lv < L else ct[b:b + 1])
        c2s.append(ctx.mod_switch(ct[n + b:n + b + 1].contiguous(), L, lv) if lv < L else ct[n + b:n + b + 1])
    flat1 = torch.cat([c.reshape(-1) for c in c1s]).contiguous()
    flat2 = torch.cat([c.reshape(-1) for c in c2s]).contiguous()
    out = ctx.mult_relin_ragged(rlk, flat1, flat2, levels)
    torch.cuda.synchronize()
    h_rlk = kb.to_host(rlk)
    o_sk = kb.to_host(sk)
    off, checked = 0, set()
    for b, lv in enumerate(levels):
        size = 2 * (lv + 1) * ctx.n
        got = out[off:off + size].reshape(1, 2, lv + 1, ctx.n)
        view = ctx.level(lv)
        # the view's own batched HomMult gives the same bits
        alone = view.mult_relin(rlk, c1s[b].contiguous(), c2s[b].contiguous())
        assert torch.equal(alone, got), (name, b, lv)
        if lv not in checked:  # the oracle once per level
            checked.add(lv)
            lc = oc.level(lv)
            want = lc.mult_relin(h_rlk, kb.to_host(c1s[b])[0], kb.to_host(c2s[b])[0])
            assert np.array_equal(kb.to_host(got)[0], want), (name, b, lv)
        dec = view.decrypt(sk, got.contiguous())
        view.decode(dec)
        want_slots = (slots[b].astype(object) * slots[n + b].astype(object)) % ctx.plain
        assert np.array_equal(kb.to_host(dec)[0].astype(object), want_slots), (name, b, lv)
        off += size
    assert off == out.numel()
    del o_sk


def test_level_view_encrypts_and_decrypts():
    import kunpbfv_amd as kb
    ctx, oc = _ctxs(SETS["k1_n8192"])
    sk = ctx.keygen_sk(bytes([50]) * 32)
    view = ctx.level(2)
    assert (view.qn, view.rn, view.dnum) == (3, 4, 3)
    pk2 = view.keygen_pk(bytes([51]) * 32, sk)
    assert tuple(pk2.shape) == (2, 3, ctx.n)
    rng = np.random.default_rng(52)
    slots = rng.integers(0, ctx.plain, size=(3, ctx.n), dtype=np.uint64)
    md = kb.to_device(slots)
    view.encode(md)
    ct = view.encrypt(bytes([53]) * 32, pk2, md)
    assert tuple(ct.shape) == (3, 2, 3, ctx.n)
    dec = view.decrypt(sk, ct)
    view.decode(dec)
    assert np.array_equal(kb.to_host(dec), slots)
    # operations that need the full level are refused on a view
    with pytest.raises(kb.KbfvError):
        view.keygen_rlk(bytes([54]) * 32, sk)


def test_ragged_levels_c4_shape():
    import torch
    import kunpbfv_amd as kb
    ctx, oc = _ctxs(P.chain60(16, 28, 1, 4, 2))
    L = ctx.qn - 1
    levels = [L, L, 13, 13]
    n = len(levels)
    sk, pk, rlk, slots, ct = _fresh(ctx, n, 60)
    c1s = [ct[b:b + 1] if lv == L else ctx.mod_switch(ct[b:b + 1].contiguous(), L, lv) for b, lv in enumerate(levels)]
    c2s = [ct[n + b:n + b + 1] if lv == L else ctx.mod_switch(ct[n + b:n + b + 1].contiguous(), L, lv)
           for b, lv in enumerate(levels)]
    flat1 = torch.cat([c.reshape(-1) for c in c1s]).contiguous()
    flat2 = torch.cat([c.reshape(-1) for c in c2s]).contiguous()
    out = ctx.mult_relin_ragged(rlk, flat1, flat2, levels)
    torch.cuda.synchronize()
    h_rlk = kb.to_host(rlk)
    off = 0
    for b, lv in enumerate(levels):
        size = 2 * (lv + 1) * ctx.n
        got = kb.to_host(out[off:off + size].reshape(2, lv + 1, ctx.n))
        if b in (1, 2):
            want = oc.level(lv).mult_relin(h_rlk, kb.to_host(c1s[b])[0], kb.to_host(c2s[b])[0])
            assert np.array_equal(got, want), (b, lv)
        off += size
