"""GPU: ragged batches (kbfv_ntt_desc, kbfv_poly_binop_desc).  Polynomials of different RNS
lengths in one launch per pass -- a full-level ciphertext polynomial over Q, a
level-0 one (one limb), a key polynomial over P+Q, an empty descriptor, a T
polynomial with a strided limb layout -- each equal bit for bit to the oracle's
NTT over its own modulus range, inverse back to the input; the call captured in
a HIP graph and replayed, with eager work between replays, gives the same bits
every time; malformed descriptors are rejected at upload; the element-wise
ragged operations match big-integer arithmetic, eager and replayed."""
import numpy as np
import pytest

import oracle_lib as orc
import params as P

pytestmark = pytest.mark.gpu

SETS = {"chain60_n8192": P.chain60(13, 3, 1, 3, 1), "ref14_k2": P.REF14}


@pytest.fixture(scope="module", params=list(SETS))
def pair(request):
    import kunpbfv_amd as kb
    d = SETS[request.param]
    ctx = kb.Context(d["log_n"], d["q"], d["p"], d["t"], d["gamma"], d["plain"])
    oc = orc.Ctx(d["log_n"], d["q"], d["p"], d["t"], d["gamma"], d["plain"])
    return kb, ctx, oc


def rand_limbs(rng, mods, n):
    return np.stack([rng.integers(0, int(q), size=n, dtype=np.uint64) for q in mods])


def build(kb, ctx, rng):
    """(items, host inputs, (mod0, limbs) per item, output tensors)"""
    N = ctx.n
    ranges = [(ctx.oQ, ctx.qn), (ctx.oQ, 1), (ctx.oP, ctx.pn + ctx.qn), (ctx.oQ, 0), (ctx.oT, ctx.tn)]
    items, xs, outs = [], [], []
    for i, (mod0, limbs) in enumerate(ranges):
        if limbs == 0:
            z = ctx.empty(1, N)
            items.append((z, z, mod0, 0))
            xs.append(None)
            outs.append(None)
            continue
        x = rand_limbs(rng, ctx.mod[mod0:mod0 + limbs], N)
        xs.append(x)
        if i == 4:  # strided source: limb l at l*2N, written densely
            src = ctx.zeros(2 * limbs, N)
            src.view(limbs, 2, N)[:, 0, :] = kb.to_device(x)
            dst = ctx.empty(limbs, N)
            items.append((src, dst, mod0, limbs, 2 * N, N))
        elif i == 1:  # in place
            dst = kb.to_device(x)
            items.append((dst, dst, mod0, limbs))
        else:
            dst = ctx.empty(limbs, N)
            items.append((kb.to_device(x), dst, mod0, limbs))
        outs.append(dst)
    return items, xs, ranges, outs


def test_ragged_forward_inverse_match_oracle(pair):
    kb, ctx, oc = pair
    rng = np.random.default_rng(31)
    items, xs, ranges, outs = build(kb, ctx, rng)
    batch = ctx.ragged(items)
    assert batch.total == sum(l for _, l in ranges)
    ctx.ntt_ragged(+1, batch)
    for x, (mod0, limbs), out in zip(xs, ranges, outs):
        if x is None:
            continue
        assert np.array_equal(kb.to_host(out), oc.ntt_fwd(x, mod0, limbs)), (mod0, limbs)
    # inverse in place on the outputs
    inv = ctx.ragged([(o, o, m0, l) for o, (m0, l) in zip(outs, ranges) if o is not None])
    ctx.ntt_ragged(-1, inv)
    for x, out in zip(xs, outs):
        if x is not None:
            assert np.array_equal(kb.to_host(out), x)


def test_ragged_graph_replays(pair):
    import torch
    kb, ctx, oc = pair
    rng = np.random.default_rng(32)
    ranges = [(ctx.oQ, ctx.qn), (ctx.oQ, 1), (ctx.oP, ctx.pn + ctx.qn)]
    xs = [rand_limbs(rng, ctx.mod[m0:m0 + l], ctx.n) for m0, l in ranges]
    srcs = [kb.to_device(x) for x in xs]
    dsts = [ctx.empty(l, ctx.n) for _, l in ranges]
    wants = [oc.ntt_fwd(x, m0, l) for x, (m0, l) in zip(xs, ranges)]
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        batch = ctx.ragged([(a, b, m0, l) for a, b, (m0, l) in zip(srcs, dsts, ranges)], stream=s)
        ctx.ntt_ragged(+1, batch, stream=s)  # warm-up outside capture
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        ctx.ntt_ragged(+1, batch)
    for rep in range(3):
        for d in dsts:
            d.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for d, w in zip(dsts, wants):
            assert np.array_equal(kb.to_host(d), w), rep


def test_upload_rejects_malformed_descriptors(pair):
    kb, ctx, _ = pair
    x = ctx.empty(ctx.qn, ctx.n)
    with pytest.raises(kb.KbfvError):
        ctx.ragged([(x, x, ctx.oQ, ctx.nmod)])           # past the modulus table
    with pytest.raises(kb.KbfvError):
        ctx.ragged([(x, x, ctx.oQ, 0)])                  # nothing to transform
    with pytest.raises(kb.KbfvError):
        ctx.ragged([(x, x, ctx.oQ, ctx.qn, 1, ctx.n)])   # overlapping source limbs
    with pytest.raises(kb.KbfvError):
        ctx.ragged([(x, x, ctx.oQ, ctx.qn + 1)])         # tensor too small


@pytest.mark.parametrize("op", [0, 1, 2])
def test_ragged_binop_matches_modular_arithmetic(pair, op):
    """kbfv_poly_binop_desc over a full-level Q polynomial, a level-0 one, a
    P+Q key polynomial (out aliasing an input) and an empty descriptor, against
    Python big-integer arithmetic mod each limb's prime; then the same launch
    replayed from a HIP graph."""
    import torch
    kb, ctx, _ = pair
    rng = np.random.default_rng(40 + op)
    ranges = [(ctx.oQ, ctx.qn), (ctx.oQ, 1), (ctx.oP, ctx.pn + ctx.qn), (ctx.oQ, 0)]
    items, wants, outs = [], [], []
    for i, (m0, l) in enumerate(ranges):
        if l == 0:
            z = ctx.empty(1, ctx.n)
            items.append((z, z, z, m0, 0))
            continue
        mods = ctx.mod[m0:m0 + l]
        a, b = rand_limbs(rng, mods, ctx.n), rand_limbs(rng, mods, ctx.n)
        qq = np.array([int(x) for x in mods], dtype=object)[:, None]
        ao, bo = a.astype(object), b.astype(object)
        want = ((ao + bo) if op == 0 else (ao - bo) if op == 1 else (ao * bo)) % qq
        ad, bd = kb.to_device(a), kb.to_device(b)
        out = ad if i == 2 else ctx.empty(l, ctx.n)
        items.append((ad, bd, out, m0, l))
        wants.append((want, a, b))
        outs.append((out, ad, bd))
    batch = ctx.ragged_binop(items)
    ctx.binop_ragged(op, batch)
    for (want, _, _), (out, _, _) in zip(wants, outs):
        assert np.array_equal(kb.to_host(out).astype(object), want)
    # graph replay on the non-aliasing descriptors only (an aliased output
    # would feed each replay's result into the next)
    g_items = [it for i, it in enumerate(items) if i in (0, 1)]
    gb = ctx.ragged_binop(g_items)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        ctx.binop_ragged(op, gb, stream=s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        ctx.binop_ragged(op, gb)
    for rep in range(3):
        for it in g_items:
            it[2].zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for k in range(2):
            assert np.array_equal(kb.to_host(g_items[k][2]).astype(object), wants[k][0]), rep
