"""The allocation-free HomMult entry point (kbfv_mult_relin_noalloc) against
kbfv_mult_relin, through the C ABI: the same bits on two calls in a row, the
second with a smaller batch (the scratch array keeps the first call's
contents), at a shape whose batch runs in several scratch-sized chunks; and a
clear rejection of a shape whose single ciphertext needs more scratch than the
array holds."""
import numpy as np
import pytest

import params as P

pytestmark = pytest.mark.gpu

SCRATCH = 128 << 20


def _ctx(log_n, qn):
    import kunpbfv_amd as kb
    d = P.chain60(log_n, qn, 1, 4, 2)
    return kb.Context(d["log_n"], d["q"], d["p"], d["t"], d["gamma"], d["plain"])


def _inputs(ctx, batch, seed):
    import kunpbfv_amd as kb
    sk = ctx.keygen_sk(bytes([seed]) * 32)
    pk = ctx.keygen_pk(bytes([seed + 1]) * 32, sk)
    rlk = ctx.keygen_rlk(bytes([seed + 2]) * 32, sk)
    rng = np.random.default_rng(seed)
    m = rng.integers(0, ctx.plain, size=(2 * batch, ctx.n), dtype=np.uint64)
    md = kb.to_device(m)
    ctx.encode(md)
    ct = ctx.encrypt(bytes([seed + 3]) * 32, pk, md)
    return rlk, ct[:batch].contiguous(), ct[batch:].contiguous()


def _noalloc(ctx, rlk, c1, c2, out):
    import kunpbfv_amd as kb
    return kb.lib.kbfv_mult_relin_noalloc(ctx.h, kb.ptr(rlk), kb.ptr(c1), kb.ptr(c2), kb.ptr(out), c1.shape[0],
                                          kb.stream_ptr())


@pytest.mark.parametrize("log_n,qn,batches", [(13, 2, (6, 2)), (15, 13, (5, 3))])
def test_noalloc_equals_workspace_entry_point(log_n, qn, batches):
    import torch
    import kunpbfv_amd as kb
    ctx = _ctx(log_n, qn)
    cap = kb.lib.kbfv_mult_noalloc_max_batch(ctx.h)
    assert cap >= 1 and ctx.mult_workspace_bytes(cap) <= SCRATCH < ctx.mult_workspace_bytes(cap + 1)
    if log_n == 15:
        assert cap < batches[1] < batches[0]  # both calls run in several chunks
    rlk, c1, c2 = _inputs(ctx, batches[0], 20 + log_n)
    for i, b in enumerate(batches):
        a1, a2 = c1[:b].contiguous(), c2[:b].contiguous()
        want = ctx.mult_relin(rlk, a1, a2)
        got = ctx.empty(*a1.shape)
        got.fill_(-1)
        kb.check(_noalloc(ctx, rlk, a1, a2, got))
        torch.cuda.synchronize()
        assert torch.equal(got, want), (log_n, i, b)


def test_noalloc_rejects_shape_beyond_scratch():
    import kunpbfv_amd as kb
    ctx = _ctx(16, 28)
    assert ctx.mult_workspace_bytes(1) > SCRATCH
    assert kb.lib.kbfv_mult_noalloc_max_batch(ctx.h) == 0
    dummy = ctx.empty(1, 2, ctx.qn, ctx.n)
    rc = _noalloc(ctx, dummy, dummy, dummy, dummy)
    assert rc == -1
    assert b"scratch" in kb.lib.kbfv_last_error()
    # nothing was launched: the context still works for the workspace entry point's sizing
    assert kb.lib.kbfv_mult_workspace_bytes(ctx.h, 1) == ctx.mult_workspace_bytes(1)
