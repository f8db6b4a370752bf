"""GPU: the batched NTT alone at its benchmark shape -- N=2^14, 8 x 60-bit
primes (1 mod 2^17, so 1 mod 2^15), 256 polynomials (2048 limbs) of seeded
uniform residues: forward bit-exact with the oracle's NTT and inverse back to
the input, under the built-in plan, the single-pass plan (one workgroup per
limb) and the plan the autotuner picks for this shape."""
import numpy as np
import pytest

import oracle_lib as orc
import params as P

pytestmark = pytest.mark.gpu


def test_c2_shape_every_kind_of_plan():
    import kunpbfv_amd as kb
    d = P.chain60(14, 8, 1, 4, 2)
    ctx = kb.Context(d["log_n"], d["q"], d["p"], d["t"], d["gamma"], d["plain"])
    oc = orc.Ctx(d["log_n"], d["q"], d["p"], d["t"], d["gamma"], d["plain"])
    mod0, limbs, polys = ctx.oQ, ctx.qn, 256
    rng = np.random.default_rng(2024)
    x = np.empty((polys, limbs, ctx.n), dtype=np.uint64)
    for i in range(limbs):
        x[:, i, :] = rng.integers(0, int(ctx.mod[mod0 + i]), size=(polys, ctx.n), dtype=np.uint64)
    want = oc.ntt_fwd(x, mod0, limbs, polys)
    bucket = kb.lib.kbfv_ntt_plan_bucket(limbs * polys)
    xd = kb.to_device(x)
    yd = ctx.empty(*x.shape)
    tuned = {}
    for direction in (+1, -1):
        tuned[direction] = ctx.autotune_ntt(direction, mod0, limbs, polys, reps=2)[0]
    for name, plans in (("built-in", (0, 0)), ("single pass", (14 | 1 << 12, 14 | 1 << 12)),
                        ("tuned", (tuned[+1], tuned[-1]))):
        ctx.set_ntt_plan(+1, bucket, plans[0])
        ctx.set_ntt_plan(-1, bucket, plans[1])
        ctx.ntt(xd, yd, +1, mod0, limbs, polys=polys)
        assert np.array_equal(kb.to_host(yd), want), name
        ctx.ntt(yd, yd, -1, mod0, limbs, polys=polys)
        assert np.array_equal(kb.to_host(yd), x), name
