"""GPU: the split-IP seam (the reference's multAndEqual_beforeIP_23 /
multAndEqual_afterIP_23 around the key-stationary inner product) gives the
fused kbfv_mult_relin result bit for bit, for batches that fill the matrix-core
inner product (17) and that do not (3), with the two halves run in chunks
(small workspace) and the inner-product output left untouched.  Each half is
also held to the oracle on its own: the digits handed to the inner product
equal the oracle's digits of a1a2 over T (modUp Qj -> T, then the NTT over
T), the parked linear terms equal the oracle's HPS-scaled a1b2 + a2b1 and
b1b2, and the final ciphertexts equal the oracle's HomMult."""
import numpy as np
import pytest

import oracle_lib as orc
import params as P

pytestmark = pytest.mark.gpu

SETS = {"chain60_n8192": P.chain60(13, 3, 1, 3, 1), "ref14_k2": P.REF14}


@pytest.fixture(scope="module", params=list(SETS))
def setup(request):
    import kunpbfv_amd as kb
    d = SETS[request.param]
    ctx = kb.Context(d["log_n"], d["q"], d["p"], d["t"], d["gamma"], d["plain"])
    sk = ctx.keygen_sk(bytes([1]) * 32)
    pk = ctx.keygen_pk(bytes([2]) * 32, sk)
    rlk = ctx.keygen_rlk(bytes([3]) * 32, sk)
    oc = orc.Ctx(d["log_n"], d["q"], d["p"], d["t"], d["gamma"], d["plain"])
    return kb, ctx, sk, pk, rlk, oc


@pytest.mark.parametrize("batch", [3, 17])
def test_split_equals_fused(setup, batch):
    import torch
    kb, ctx, sk, pk, rlk, oc = setup
    rng = np.random.default_rng(batch)
    m = kb.to_device(rng.integers(0, ctx.plain, size=(2 * batch, ctx.n), dtype=np.uint64))
    ctx.encode(m)
    ct = ctx.encrypt(bytes([4]) * 32, pk, m)
    c1, c2 = ct[:batch].contiguous(), ct[batch:].contiguous()
    fused = ctx.mult_relin(rlk, c1, c2)
    ip_in, d01 = ctx.mult_before_ip(c1, c2)
    ip_out = ctx.ip_batch(ip_in, rlk)
    keep = ip_out.clone()
    out = ctx.mult_after_ip(ip_out, d01)
    torch.cuda.synchronize()
    assert torch.equal(out, fused)
    assert torch.equal(ip_out, keep)
    # chunked: a workspace for one ciphertext only
    ws = torch.empty(ctx.mult_workspace_bytes(1), dtype=torch.uint8, device="cuda")
    ip_in2, d012 = ctx.mult_before_ip(c1, c2, ws=ws)
    out2 = ctx.mult_after_ip(ctx.ip_batch(ip_in2, rlk), d012, ws=ws)
    torch.cuda.synchronize()
    assert torch.equal(ip_in2, ip_in) and torch.equal(d012, d01) and torch.equal(out2, fused)
    # each half against the oracle, on the first and the last pair
    h1, h2, o_rlk = kb.to_host(c1), kb.to_host(c2), kb.to_host(rlk)
    g_ip, g_d01, g_out = kb.to_host(ip_in), kb.to_host(d01), kb.to_host(out)
    for b in (0, batch - 1):
        d = oc.mult_tensor(h1[b], h2[b])  # coefficient form over Q: a1a2, a1b2 + a2b1, b1b2
        digits = oc.ntt_fwd(oc.modup_qj_to_t(d[0]), oc.oT, oc.tn, polys=oc.dnum)
        assert np.array_equal(g_ip[b], digits), b
        assert np.array_equal(g_d01[b], d[1:]), b
        assert np.array_equal(g_out[b], oc.mult_relin(o_rlk, h1[b], h2[b])), b
