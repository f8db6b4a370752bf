"""GPU parity of the key-switch inner product on batches (the int8 matrix-core
path takes batches of 16 and more, in tiles of 32 ciphertexts): bit-exact with
the oracle per ciphertext, for ragged batches and keys of 2 to 28 digits."""
import numpy as np
import pytest

import oracle_lib as orc
import params as P

pytestmark = pytest.mark.gpu

SETS = {
    "ref14_k2": dict(P.REF14, log_n=12),                  # dnum 2, 3 key blocks
    "ref15_chain": dict(P.REF15, log_n=12),               # dnum 15, 8 key blocks
    "chain60_28": P.chain60(12, 28, 1, 4, 2),             # dnum 28, 15 key blocks
}


@pytest.fixture(scope="module", params=list(SETS))
def pair(request):
    import kunpbfv_amd as kb
    d = SETS[request.param]
    ctx = kb.Context(d["log_n"], d["q"], d["p"], d["t"], d["gamma"], d["plain"])
    oc = orc.Ctx(d["log_n"], d["q"], d["p"], d["t"], d["gamma"], d["plain"])
    return ctx, oc


@pytest.mark.parametrize("batch", [16, 33])
def test_inner_product_batch(pair, batch):
    import kunpbfv_amd as kb
    ctx, oc = pair
    rng = np.random.default_rng(batch * 7 + ctx.dnum)
    N, D, B, tn = ctx.n, ctx.dnum, ctx.ri_blocks, ctx.tn
    tm = [int(v) for v in ctx.mod[ctx.oT:ctx.oT + tn]]
    cin = np.empty((batch, D, tn, N), dtype=np.uint64)
    key = np.empty((2, D, B, tn, N), dtype=np.uint64)
    for k, q in enumerate(tm):
        cin[:, :, k, :] = rng.integers(0, q, size=(batch, D, N), dtype=np.uint64)
        key[:, :, :, k, :] = rng.integers(0, q, size=(2, D, B, N), dtype=np.uint64)
        # the largest residues in the first ciphertext and the first coefficients
        cin[0, :, k, :] = q - 1
        cin[:, :, k, :16] = q - 1
        key[:, :, :, k, :8] = q - 1
    out = ctx.empty(batch, 2, B, tn, N)
    ctx.inner_product(kb.to_device(cin), kb.to_device(key), out, batch=batch)
    got = kb.to_host(out)
    for ct in range(batch):
        assert np.array_equal(got[ct], oc.inner_product(cin[ct], key)), ct
