"""GPU: a chain whose auxiliary T primes are much smaller than its Q primes
(60-bit Q, 45-bit T).  With one special prime the digit -> T conversion can
read a Q limb straight into the forward NTT over T only when q < 8 t for every
pair (the NTT's lazy input range); here it cannot, so the context takes the
explicit modUp Q_j -> T kernel instead.  HomMult + relinearisation equals the
oracle bit for bit and decrypts to the slot-wise product."""
import numpy as np
import pytest

import oracle_lib as orc

pytestmark = pytest.mark.gpu


def test_small_t_primes_hommult_matches_oracle():
    import kunpbfv_amd as kb
    q, p, _t = orc.chain_60bit(3, 1, 1)
    t = orc.primes_descending(1 << 45, 8192, 6)
    assert max(q) >= 8 * min(t)
    ctx = kb.Context(12, q, p, t, 2, 65537)
    oc = orc.Ctx(12, q, p, t, 2, 65537)
    sk = ctx.keygen_sk(bytes([1]) * 32)
    pk = ctx.keygen_pk(bytes([2]) * 32, sk)
    rlk = ctx.keygen_rlk(bytes([3]) * 32, sk)
    o_rlk = oc.keygen_rlk(bytes([3]) * 32, oc.keygen_sk(bytes([1]) * 32))
    assert np.array_equal(kb.to_host(rlk), o_rlk)
    rng = np.random.default_rng(9)
    slots = rng.integers(0, 65537, size=(4, ctx.n), dtype=np.uint64)
    md = kb.to_device(slots)
    ctx.encode(md)
    ct = ctx.encrypt(bytes([4]) * 32, pk, md)
    out = kb.to_host(ctx.mult_relin(rlk, ct[:2].contiguous(), ct[2:].contiguous()))
    h = kb.to_host(ct)
    for b in range(2):
        assert np.array_equal(out[b], oc.mult_relin(o_rlk, h[b], h[2 + b])), b
    dec = ctx.decrypt(sk, kb.to_device(out))
    ctx.decode(dec)
    want = (slots[:2].astype(object) * slots[2:].astype(object)) % 65537
    assert np.array_equal(kb.to_host(dec).astype(object), want)
