"""GPU parity at N = 2^16 with 28 + 1 60-bit primes (4 T primes, gamma 2) for
the two kernels that only switch on at that shape's full batch sizes:

* the RGSW x RLWE external product with gadget base 2^20 (168 gadget rows,
  k_ntt_digit_col<16, ...> fusing the digit split into the column pass, the
  gadget MAC): the RGSW ciphertext and two products of a batch of 3 equal the
  oracle bit for bit, and decrypt to the shifted message;
* the int8-MFMA key-switch inner product (k_ip_mfma, batch >= 16): three
  entries of a batch of 16 equal the oracle.

The oracle needs about 30 s for the RGSW ciphertext and 15 s per product at
this size; the two products run on two host threads."""
import concurrent.futures as cf

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
    assert ctx.n == 65536 and ctx.qn == 28 and ctx.dnum == 28 and ctx.ri_blocks == 15
    return ctx, oc


def test_external_product_n65536_base20_matches_oracle(pair):
    import kunpbfv_amd as kb
    ctx, oc = pair
    base = 20
    assert ctx.rgsw_rows(base) == 168
    s_sk, s_pk, s_rg, s_enc = (bytes([i]) * 32 for i in (41, 42, 43, 44))
    sk = ctx.keygen_sk(s_sk)
    pk = ctx.keygen_pk(s_pk, sk)
    mg = np.zeros(ctx.n, dtype=np.uint64)
    mg[3] = 1  # RGSW(X^3): the product is the RLWE message shifted by 3
    rgsw = ctx.rgsw_encrypt(s_rg, sk, kb.to_device(mg[None])[0], base)
    o_sk = oc.keygen_sk(s_sk)
    o_rgsw = oc.rgsw_encrypt(s_rg, o_sk, mg, base)
    assert np.array_equal(kb.to_host(rgsw), o_rgsw)

    rng = np.random.default_rng(66)
    m = rng.integers(0, ctx.plain, size=(3, ctx.n), dtype=np.uint64)
    md = kb.to_device(m)
    ctx.encode(md)
    ct = ctx.encrypt(s_enc, pk, md)
    # entry 2: the largest residue in every coefficient (every digit 2^20 - 1 or the top digit)
    ct[2] = kb.to_device(np.broadcast_to((ctx.mod[ctx.oQ:ctx.oQ + ctx.qn] - 1)[None, :, None],
                                         (2, ctx.qn, ctx.n)).copy())
    out = kb.to_host(ctx.external_product(rgsw, ct, base))
    cth = kb.to_host(ct)
    with cf.ThreadPoolExecutor(2) as ex:
        want = list(ex.map(lambda b: oc.external_product(o_rgsw, cth[b], base), (0, 2)))
    assert np.array_equal(out[0], want[0])
    assert np.array_equal(out[2], want[1])

    dec = ctx.decrypt(sk, kb.to_device(out[:2].copy()))
    ctx.decode(dec)
    got = kb.to_host(dec).astype(np.int64)
    enc = kb.to_host(md).astype(np.int64)
    t = ctx.plain
    for b in range(2):
        want_b = np.concatenate([-enc[b, -3:], enc[b, :-3]]) % t
        assert np.array_equal(got[b] % t, want_b), b


def test_inner_product_mfma_n65536_batch16_matches_oracle(pair):
    import kunpbfv_amd as kb
    ctx, oc = pair
    batch = 16  # the int8-MFMA path's threshold
    N, D, B, tn = ctx.n, ctx.dnum, ctx.ri_blocks, ctx.tn
    tm = [int(v) for v in ctx.mod[ctx.oT:ctx.oT + tn]]
    rng = np.random.default_rng(67)
    cin = np.empty((batch, D, tn, N), dtype=np.uint64)
    key = np.empty((2, D, B, tn, N), dtype=np.uint64)
    for k, q in enumerate(tm):
        cin[:, :, k, :] = rng.integers(0, q, size=(batch, D, N), dtype=np.uint64)
        key[:, :, :, k, :] = rng.integers(0, q, size=(2, D, B, N), dtype=np.uint64)
        cin[batch - 1, :, k, :] = q - 1  # extreme residues in the last entry
        key[:, :, :, k, :64] = q - 1
    out = ctx.empty(batch, 2, B, tn, N)
    ctx.inner_product(kb.to_device(cin), kb.to_device(key), out, batch=batch)
    got = kb.to_host(out)
    with cf.ThreadPoolExecutor(3) as ex:
        want = list(ex.map(lambda b: oc.inner_product(cin[b], key), (0, 7, batch - 1)))
    for b, w in zip((0, 7, batch - 1), want):
        assert np.array_equal(got[b], w), b
