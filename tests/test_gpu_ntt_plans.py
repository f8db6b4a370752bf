"""GPU: NTT plans and the autotuner.  Every plan of the plan space (split point,
tile of each pass, grid order of each pass) computes the oracle's forward NTT
and inverts it exactly, at every ring degree; HomMult+relin (which also runs the
unscaled inverse) gives the same bits under non-default plans; the autotuner
stores a valid plan for the bucket it tuned and the transform stays bit-exact."""
import numpy as np
import pytest

import oracle_lib as orc
import params as P

pytestmark = pytest.mark.gpu


def make(log_n):
    import kunpbfv_amd as kb
    d = P.chain60(log_n, 3, 1, 3, 1)
    ctx = kb.Context(d["log_n"], d["q"], d["p"], d["t"], d["gamma"], d["plain"])
    oc = orc.Ctx(d["log_n"], d["q"], d["p"], d["t"], d["gamma"], d["plain"])
    return kb, ctx, oc


def rand_limbs(rng, mods, n, polys):
    out = np.empty((polys, len(mods), n), dtype=np.uint64)
    for i, q in enumerate(mods):
        out[:, i, :] = rng.integers(0, int(q), size=(polys, n), dtype=np.uint64)
    return out


def set_all(kb, ctx, plan):
    for b in range(kb.lib.kbfv_ntt_plan_bucket(1 << 40) + 1):
        ctx.set_ntt_plan(+1, b, plan)
        ctx.set_ntt_plan(-1, b, plan)


@pytest.mark.parametrize("log_n", [12, 13, 14, 15, 16])
def test_every_plan_matches_oracle(log_n):
    kb, ctx, oc = make(log_n)
    rng = np.random.default_rng(100 + log_n)
    mod0, limbs, polys = ctx.oQ, ctx.qn, 2
    x = rand_limbs(rng, ctx.mod[mod0:mod0 + limbs], ctx.n, polys)
    want = oc.ntt_fwd(x, mod0, limbs, polys)
    xd = kb.to_device(x)
    yd = ctx.empty(*x.shape)
    plans = kb.ntt_plan_candidates(log_n)
    assert len(plans) >= 20
    bad = []
    for p in plans:
        set_all(kb, ctx, p)
        ctx.ntt(xd, yd, +1, mod0, limbs, polys=polys)
        if not np.array_equal(kb.to_host(yd), want):
            bad.append(("forward", hex(p)))
        ctx.ntt(yd, yd, -1, mod0, limbs, polys=polys)
        if not np.array_equal(kb.to_host(yd), x):
            bad.append(("inverse", hex(p)))
    assert not bad, bad


@pytest.mark.parametrize("log_n", [13, 14])
def test_hommult_same_bits_under_other_plans(log_n):
    kb, ctx, _ = make(log_n)
    sk = ctx.keygen_sk(bytes([61]) * 32)
    pk = ctx.keygen_pk(bytes([62]) * 32, sk)
    rlk = ctx.keygen_rlk(bytes([63]) * 32, sk)
    rng = np.random.default_rng(64)
    m = rng.integers(0, ctx.plain, size=(2, ctx.n), dtype=np.uint64)
    ct = ctx.encrypt(bytes([65]) * 32, pk, kb.to_device(m))
    c1, c2 = ct[0:1].contiguous(), ct[1:2].contiguous()
    base = kb.to_host(ctx.mult_relin(rlk, c1, c2))
    plans = kb.ntt_plan_candidates(log_n)
    # every split and tile of the column pass (the unscaled inverse's last pass)
    # and a spread of the rest
    pick = {}
    for p in plans:
        pick.setdefault((p & 0xFF), p)
    chosen = sorted(set(pick.values()) | set(plans[::7]))
    for p in chosen:
        set_all(kb, ctx, p)
        got = kb.to_host(ctx.mult_relin(rlk, c1, c2))
        assert np.array_equal(got, base), hex(p)


def test_autotune_stores_a_valid_plan_and_stays_exact():
    import torch
    kb, ctx, oc = make(15)
    mod0, limbs, polys = ctx.oQ, ctx.qn, 40
    scratch = torch.empty(polys * limbs * ctx.n, dtype=torch.int64, device="cuda")
    for direction in (+1, -1):
        plan, ms, bucket = ctx.autotune_ntt(direction, mod0, limbs, polys, reps=2, scratch=scratch)
        assert plan != 0 and kb.lib.kbfv_ntt_plan_valid(15, plan) == 0
        assert ms > 0 and bucket == kb.lib.kbfv_ntt_plan_bucket(limbs * polys)
        assert ctx.ntt_plan(direction, bucket) == plan
    rng = np.random.default_rng(7)
    x = rand_limbs(rng, ctx.mod[mod0:mod0 + limbs], ctx.n, polys)
    xd = kb.to_device(x)
    ctx.ntt(xd, xd, +1, mod0, limbs, polys=polys)
    assert np.array_equal(kb.to_host(xd), oc.ntt_fwd(x, mod0, limbs, polys))
    ctx.ntt(xd, xd, -1, mod0, limbs, polys=polys)
    assert np.array_equal(kb.to_host(xd), x)


def test_plan_set_rejects_invalid_words():
    kb, ctx, _ = make(14)
    with pytest.raises(kb.KbfvError):
        ctx.set_ntt_plan(+1, 0, 0x1234567)
    with pytest.raises(kb.KbfvError):
        ctx.set_ntt_plan(+1, 9, 0)


def test_autotune_json_cache(tmp_path):
    """A tuned context writes its plans to the JSON cache; a second context of
    the same parameters takes them from it without timing, a malformed entry is
    tuned again and replaced, and a different batch is a separate entry."""
    import json
    kb, ctx, _ = make(14)
    path = str(tmp_path / "plans.json")
    first = ctx.autotune(4, reps=1, cache=path)
    data = json.load(open(path))
    assert len(data) == 1
    kb2, ctx2, _ = make(14)
    second = ctx2.autotune(4, reps=1, cache=path)
    assert second == first
    for (d, b), (plan, _, _) in first.items():
        assert ctx2.ntt_plan(d, b) == plan
    key = next(iter(data))
    data[key][0][2] = 1 << 20  # not a plan word
    json.dump(data, open(path, "w"))
    third = ctx2.autotune(4, reps=1, cache=path)
    assert all(kb.lib.kbfv_ntt_plan_valid(14, p) == 0 for p, _, _ in third.values())
    assert json.load(open(path))[key][0][2] != 1 << 20
    ctx2.autotune(8, reps=1, cache=path)
    assert len(json.load(open(path))) == 2


@pytest.mark.parametrize("layout", ["polys", "batch"])
def test_chunked_plans_match_oracle(layout):
    """Chunked two-pass plans (both passes per chunk of 32 MB) over a 36 MB
    launch at N = 2^16, chunked along the batch or the polynomials: the
    oracle's forward NTT and the exact inverse, with a short last chunk."""
    kb, ctx, oc = make(16)
    rng = np.random.default_rng(300)
    mod0, limbs, n = ctx.oQ, ctx.qn, 24
    x = rand_limbs(rng, ctx.mod[mod0:mod0 + limbs], ctx.n, n)  # [24][3][N] = 36 MiB
    want = oc.ntt_fwd(x, mod0, limbs, n)
    N = ctx.n
    for p in (0x2cc8, 0x3cd8 | 1 << 12):
        plan = (p & 0x3fff) | 2 << 14
        assert kb.lib.kbfv_ntt_plan_valid(16, plan) == 0
        set_all(kb, ctx, plan)
        xd = kb.to_device(x)
        if layout == "polys":
            ctx.ntt(xd, xd, +1, mod0, limbs, polys=n)
            assert np.array_equal(kb.to_host(xd), want), hex(plan)
            ctx.ntt(xd, xd, -1, mod0, limbs, polys=n)
        else:
            s = (limbs * N, 0, N)
            ctx.ntt(xd, xd, +1, mod0, limbs, polys=1, batch=n, s=s)
            assert np.array_equal(kb.to_host(xd), want), hex(plan)
            ctx.ntt(xd, xd, -1, mod0, limbs, polys=1, batch=n, s=s)
        assert np.array_equal(kb.to_host(xd), x), hex(plan)
