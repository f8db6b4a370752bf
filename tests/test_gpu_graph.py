"""GPU: the HIP-graph form of the batched HomMult+relin.  The graph is replayed
several times with eager work and synchronisations between replays, new
ciphertexts are copied into the captured input buffers between replays, and
every replay must equal the eager call on the same inputs bit for bit (and the
eager call must equal the oracle).  Batches of 3 (scalar inner product) and 20
(matrix-core inner product) are covered."""
import numpy as np
import pytest

import oracle_lib as orc
import params as P

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup():
    import kunpbfv_amd as kb
    d = P.chain60(13, 3, 1, 3, 1)
    ctx = kb.Context(d["log_n"], d["q"], d["p"], d["t"], d["gamma"], d["plain"])
    oc = orc.Ctx(d["log_n"], d["q"], d["p"], d["t"], d["gamma"], d["plain"])
    sk = ctx.keygen_sk(bytes([1]) * 32)
    pk = ctx.keygen_pk(bytes([2]) * 32, sk)
    rlk = ctx.keygen_rlk(bytes([3]) * 32, sk)
    return kb, ctx, oc, sk, pk, rlk


def _encrypt(kb, ctx, pk, seed, batch, rng):
    m = rng.integers(0, ctx.plain, size=(batch, ctx.n), dtype=np.uint64)
    md = kb.to_device(m)
    ctx.encode(md)
    return m, ctx.encrypt(bytes([seed]) * 32, pk, md)


@pytest.mark.parametrize("batch", [3, 20])
def test_graph_replays_match_eager(setup, batch):
    import torch
    kb, ctx, oc, sk, pk, rlk = setup
    rng = np.random.default_rng(batch)
    _, c1 = _encrypt(kb, ctx, pk, 10, batch, rng)
    _, c2 = _encrypt(kb, ctx, pk, 11, batch, rng)
    nb = ctx.mult_workspace_bytes(batch)
    ws = torch.empty(nb, dtype=torch.uint8, device="cuda")  # the graph's own workspace
    out_g = ctx.empty(*c1.shape)
    graph = ctx.mult_relin_graph(rlk, c1, c2, out_g, ws)
    ws_e = torch.empty(nb, dtype=torch.uint8, device="cuda")
    for rep in range(4):
        if rep:
            # new inputs in the captured buffers
            m1, n1 = _encrypt(kb, ctx, pk, 20 + rep, batch, rng)
            m2, n2 = _encrypt(kb, ctx, pk, 40 + rep, batch, rng)
            c1.copy_(n1)
            c2.copy_(n2)
        want = ctx.mult_relin(rlk, c1, c2, ws=ws_e)  # eager, other workspace
        torch.cuda.synchronize()
        out_g.zero_()
        graph.launch()
        torch.cuda.synchronize()
        assert torch.equal(out_g, want), rep
        if rep == 1:
            h1, h2, hw = kb.to_host(c1), kb.to_host(c2), kb.to_host(want)
            o_rlk = oc.keygen_rlk(bytes([3]) * 32, kb.to_host(sk))
            assert np.array_equal(hw[0], oc.mult_relin(o_rlk, h1[0], h2[0]))
            dec = ctx.decrypt(sk, out_g)
            ctx.decode(dec)
            got = kb.to_host(dec)
            want_m = (m1.astype(object) * m2.astype(object)) % ctx.plain
            assert np.array_equal(got.astype(object), want_m)
    del graph
