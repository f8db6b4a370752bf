"""GPU parity at the HIP-graph configuration's shape: N = 2^15, 13 + 1 60-bit
primes, 4 T primes, gamma = 2, t = 65537 (1 mod 2^16, so batching).  A batch
of 20 pairs takes the int8-MFMA inner product (batches of 16 and more) and the
2^15 instantiations of every pipeline kernel; two of the products are checked
bit for bit against the CPU oracle, all of them by decryption (slot-wise
products), and the captured graph replays give the eager bits."""
import numpy as np
import pytest

import oracle_lib as orc
import params as P

pytestmark = pytest.mark.gpu


def test_hommult_n32768_batch20_matches_oracle():
    import torch
    import kunpbfv_amd as kb
    d = P.chain60(15, 13, 1, 4, 2)
    ctx = kb.Context(d["log_n"], d["q"], d["p"], d["t"], d["gamma"], d["plain"])
    oc = orc.Ctx(d["log_n"], d["q"], d["p"], d["t"], d["gamma"], d["plain"])
    assert ctx.n == 32768 and ctx.qn == 13 and ctx.batching
    s = [bytes([i]) * 32 for i in (51, 52, 53)]
    sk = ctx.keygen_sk(s[0])
    pk = ctx.keygen_pk(s[1], sk)
    rlk = ctx.keygen_rlk(s[2], sk)
    o_sk = oc.keygen_sk(s[0])
    o_rlk = oc.keygen_rlk(s[2], o_sk)
    assert np.array_equal(kb.to_host(rlk), o_rlk)

    B = 20
    rng = np.random.default_rng(54)
    m1 = rng.integers(0, ctx.plain, size=(B, ctx.n), dtype=np.uint64)
    m2 = rng.integers(0, ctx.plain, size=(B, ctx.n), dtype=np.uint64)
    cts = []
    for seed, m in ((55, m1), (56, m2)):
        md = kb.to_device(m)
        ctx.encode(md)
        cts.append(ctx.encrypt(bytes([seed]) * 32, pk, md))
    c1, c2 = cts
    out = ctx.mult_relin(rlk, c1, c2)
    oh = kb.to_host(out)
    c1h, c2h = kb.to_host(c1), kb.to_host(c2)
    for b in (0, B - 1):
        assert np.array_equal(oh[b], oc.mult_relin(o_rlk, c1h[b], c2h[b])), b

    dec = ctx.decrypt(sk, out)
    ctx.decode(dec)
    want = (m1.astype(object) * m2.astype(object)) % ctx.plain
    assert np.array_equal(kb.to_host(dec).astype(object), want)

    ws = torch.empty(ctx.mult_workspace_bytes(B), dtype=torch.uint8, device="cuda")
    out_g = ctx.empty(*c1.shape)
    graph = ctx.mult_relin_graph(rlk, c1, c2, out_g, ws)
    stream = torch.cuda.current_stream()
    for rep in range(2):
        out_g.zero_()
        torch.cuda.synchronize()
        graph.launch(stream)
        torch.cuda.synchronize()
        assert torch.equal(out_g, out), rep
    del graph
