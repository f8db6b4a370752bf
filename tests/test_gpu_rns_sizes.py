"""GPU parity of the HPS base conversion and scale-and-round over basis sizes
(1 to 31 ciphertext primes of 60 bits at N=4096), several polynomials per call,
and extreme residues (0, 1, q-1): the int8 matrix-core path for up to 30
primes and the scalar path beyond, both bit-exact with the oracle."""
import numpy as np
import pytest

import oracle_lib as orc
import params as P

pytestmark = pytest.mark.gpu

QNS = [1, 2, 7, 8, 15, 16, 28, 30, 31]


def rand_mod(rng, mods, n, polys, extreme):
    out = np.empty((polys, len(mods), n), dtype=np.uint64)
    for i, q in enumerate(mods):
        q = int(q)
        out[:, i, :] = rng.integers(0, q, size=(polys, n), dtype=np.uint64)
        if extreme:
            if i % 2 == 0:  # even limbs only 0, 1 or q-1
                edge = np.array([0, 1, q - 1], dtype=np.uint64)
                out[:, i, :] = edge[rng.integers(0, 3, size=(polys, n))]
            out[:, i, : n // 8] = q - 1  # every limb q-1: the largest sums
    return out


@pytest.fixture(scope="module", params=QNS)
def pair(request):
    import kunpbfv_amd as kb
    d = P.chain60(12, request.param, 1, 4, 2)
    ctx = kb.Context(d["log_n"], d["q"], d["p"], d["t"], d["gamma"], d["plain"])
    oc = orc.Ctx(d["log_n"], d["q"], d["p"], d["t"], d["gamma"], d["plain"])
    return ctx, oc


@pytest.mark.parametrize("extreme", [False, True])
def test_bconv_sizes(pair, extreme):
    import kunpbfv_amd as kb
    ctx, oc = pair
    rng = np.random.default_rng(100 + ctx.qn)
    polys = 3
    x = rand_mod(rng, ctx.mod[ctx.oQ:ctx.oQ + ctx.qn], ctx.n, polys, extreme)
    y = ctx.empty(polys, ctx.rn, ctx.n)
    ctx.bconv(kb.to_device(x), y, True, polys=polys)
    got = kb.to_host(y)
    for p in range(polys):
        assert np.array_equal(got[p], oc.bconv_q_to_r(x[p])), p
    xr = rand_mod(rng, ctx.mod[ctx.oR:ctx.oR + ctx.rn], ctx.n, polys, extreme)
    yq = ctx.empty(polys, ctx.qn, ctx.n)
    ctx.bconv(kb.to_device(xr), yq, False, polys=polys)
    got = kb.to_host(yq)
    for p in range(polys):
        assert np.array_equal(got[p], oc.bconv_r_to_q(xr[p])), p


@pytest.mark.parametrize("extreme", [False, True])
def test_scale_round_sizes(pair, extreme):
    import kunpbfv_amd as kb
    ctx, oc = pair
    rng = np.random.default_rng(200 + ctx.qn)
    polys = 2
    sn = ctx.qn + ctx.rn
    x = rand_mod(rng, ctx.mod[ctx.oQ:ctx.oQ + sn], ctx.n, polys, extreme)
    y = ctx.empty(polys, ctx.rn, ctx.n)
    ctx.scale_round(kb.to_device(x), y, polys=polys)
    got = kb.to_host(y)
    for p in range(polys):
        assert np.array_equal(got[p], oc.scale_round(x[p])), p
