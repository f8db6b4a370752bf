"""GPU: a workspace smaller than the batch needs splits HomMult into chunks,
and the last chunk may be shorter than the others.  Every chunk must lay out
its buffers for its own size (the pipeline takes the second input's
polynomials as following the first's directly), so results with a workspace
for two ciphertexts on batches of 3 and 5 (chunks 2+1 and 2+2+1) must equal the
unchunked ones bit for bit, for kbfv_mult_relin, the split-IP seam and
kbfv_rotate."""
import numpy as np
import pytest

import params as P

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=["chain60_n8192", "ref14_k2"])
def setup(request):
    import kunpbfv_amd as kb
    d = {"chain60_n8192": P.chain60(13, 3, 1, 3, 1), "ref14_k2": P.REF14}[request.param]
    ctx = kb.Context(d["log_n"], d["q"], d["p"], d["t"], d["gamma"], d["plain"])
    sk = ctx.keygen_sk(bytes([41]) * 32)
    pk = ctx.keygen_pk(bytes([42]) * 32, sk)
    rlk = ctx.keygen_rlk(bytes([43]) * 32, sk)
    rng = np.random.default_rng(44)
    m = rng.integers(0, ctx.plain, size=(5, ctx.n), dtype=np.uint64)
    md = kb.to_device(m)
    ctx.encode(md)
    ct = ctx.encrypt(bytes([45]) * 32, pk, md)
    return kb, ctx, rlk, ct


@pytest.mark.parametrize("batch", [3, 5])
def test_short_last_chunk_mult_relin(setup, batch):
    import torch
    kb, ctx, rlk, ct = setup
    c1 = ct[:batch].contiguous()
    c2 = ct.roll(1, 0)[:batch].contiguous()
    full = kb.to_host(ctx.mult_relin(rlk, c1, c2))
    ws = torch.empty(ctx.mult_workspace_bytes(2), dtype=torch.uint8, device="cuda")
    chunked = kb.to_host(ctx.mult_relin(rlk, c1, c2, ws=ws))
    assert np.array_equal(chunked, full)


@pytest.mark.parametrize("batch", [3, 5])
def test_short_last_chunk_split_ip(setup, batch):
    import torch
    kb, ctx, rlk, ct = setup
    c1 = ct[:batch].contiguous()
    c2 = ct.roll(2, 0)[:batch].contiguous()
    full = kb.to_host(ctx.mult_relin(rlk, c1, c2))
    ws = torch.empty(ctx.mult_workspace_bytes(2), dtype=torch.uint8, device="cuda")
    ip_in, d01 = ctx.mult_before_ip(c1, c2, ws=ws)
    ip_out = ctx.ip_batch(ip_in, rlk)
    out = ctx.mult_after_ip(ip_out, d01, ws=ws)
    assert np.array_equal(kb.to_host(out), full)


@pytest.mark.parametrize("batch", [3, 5])
def test_short_last_chunk_rotate(setup, batch):
    import torch
    kb, ctx, rlk, ct = setup
    sk = ctx.keygen_sk(bytes([41]) * 32)
    g = ctx.galois_elt(1)
    gk = ctx.keygen_galois(bytes([46]) * 32, sk, g)
    c1 = ct[:batch].contiguous()
    full = kb.to_host(ctx.rotate(gk, g, c1))
    out = ctx.empty(*c1.shape)
    ws = torch.empty(kb.lib.kbfv_rotate_workspace_bytes(ctx.h, 2), dtype=torch.uint8, device="cuda")
    kb.check(kb.lib.kbfv_rotate(ctx.h, kb.ptr(gk), g, kb.ptr(c1), kb.ptr(out), batch, kb.ptr(ws), ws.numel(),
                                kb.stream_ptr(None)))
    assert np.array_equal(kb.to_host(out), full)
