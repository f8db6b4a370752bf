"""GPU: the continuous-batching request loop.  Every request served inside a
mixed batch must equal the same request run alone (batch 1): bit-exact for
HomMult+relin, external product and addition; encryptions must decrypt to their
message.  A burst of requests must actually be batched, and the serial plain
loop must give the same outputs."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rings():
    from kunpbfv_amd import serving
    return {12: serving.Ring(12, qn=2, pool=6, max_batch=16), 13: serving.Ring(13, pool=6, max_batch=16)}


def _check_alone(rings, reqs):
    import torch
    import kunpbfv_amd as kb
    from kunpbfv_amd import serving
    torch.cuda.synchronize()
    for r in reqs:
        ring = rings[r.log_n]
        if r.op == "encrypt":
            dec = ring.ctx.decrypt(ring.sk, r.out[None].contiguous())
            ring.ctx.decode(dec)
            assert np.array_equal(kb.to_host(dec)[0], ring.msgs[r.a]), (r.rid, r.op)
            continue
        alone = serving.copy_requests([r])
        ring.run(r.op, alone, stream=torch.cuda.current_stream(), keep=True)
        torch.cuda.synchronize()
        assert torch.equal(alone[0].out, r.out), (r.rid, r.op, r.log_n)


def test_poisson_mix_matches_alone(rings):
    from kunpbfv_amd import serving
    reqs = serving.poisson_requests(qps=3000, duration_s=0.04, log_ns=[12, 13], seed=3, pool=6)
    assert len(reqs) > 40 and {r.op for r in reqs} == set(serving.OPS)
    stats = serving.serve(rings, reqs, keep=True)
    assert stats["requests"] == len(reqs) and stats["p99_ms"] >= stats["p50_ms"] > 0
    _check_alone(rings, reqs)


def test_burst_is_batched_and_serial_loop_agrees(rings):
    import torch
    from kunpbfv_amd import serving
    reqs = serving.poisson_requests(qps=1e6, duration_s=1e-4, log_ns=[12, 13], seed=4, pool=6)
    assert len(reqs) >= 60
    stats = serving.serve(rings, reqs, keep=True)
    assert stats["mean_batch"] > 2, stats
    _check_alone(rings, reqs)
    ser = serving.copy_requests(reqs)
    sstats = serving.serial_plain(rings, ser, keep=True)
    assert sstats["requests"] == len(reqs)
    for a, b in zip(reqs, ser):
        if a.op != "encrypt":
            assert torch.equal(a.out, b.out), (a.rid, a.op)


def test_external_product_by_x_shifts_the_message(rings):
    import torch
    import kunpbfv_amd as kb
    from kunpbfv_amd import serving
    ring = rings[13]
    r = serving.Request(0, 13, "extprod", 0.0, 2, 0)
    # on the current stream: the decryption below reads the output there
    ring.run("extprod", [r], stream=torch.cuda.current_stream(), keep=True)
    dec = ring.ctx.decrypt(ring.sk, r.out[None].contiguous())
    got = kb.to_host(dec)[0].astype(np.int64)
    m = ring.msgs[2].astype(np.int64)
    # coefficient form of the encoded message, multiplied by X (negacyclic)
    enc = kb.to_host(ring.pt[2:3].contiguous())[0].astype(np.int64)
    want = np.empty_like(enc)
    want[1:] = enc[:-1]
    want[0] = -enc[-1]
    assert np.array_equal(got % ring.ctx.plain, want % ring.ctx.plain)
    assert m.shape == got.shape


def test_large_rings_mix_matches_alone():
    """The rings N = 2^14, 2^15 and 2^16 with their 128-bit chains (6, 13 and
    28 + 1 primes): a seeded mix, every result equal to the same request alone."""
    from kunpbfv_amd import serving
    big = {ln: serving.Ring(ln, pool=3, max_batch=8) for ln in (14, 15, 16)}
    reqs = serving.poisson_requests(qps=2000, duration_s=0.02, log_ns=[14, 15, 16], seed=5, pool=3)
    assert len(reqs) >= 20 and {r.log_n for r in reqs} == {14, 15, 16}
    stats = serving.serve(big, reqs, keep=True)
    assert stats["requests"] == len(reqs)
    _check_alone(big, reqs)


def test_admission_defers_over_budget_burst(rings):
    """A burst against a budget that holds only a few requests' working memory
    at a time: batches are cut down and deferred until running ones release
    their memory, the reserved bytes never exceed the budget, and every result
    still equals the same request alone."""
    from kunpbfv_amd import serving
    cap = max(rings[ln].batch_bytes(op, 3) for ln in rings for op in serving.OPS)
    budget = serving.MemoryBudget(capacity=cap)
    reqs = serving.poisson_requests(qps=1e6, duration_s=1e-4, log_ns=[12, 13], seed=6, pool=6)
    assert len(reqs) >= 60
    stats = serving.serve(rings, reqs, keep=True, budget=budget)
    adm = stats["admission"]
    assert stats["requests"] == len(reqs) and adm["rejected"] == 0
    assert adm["halvings"] > 0 and adm["deferrals"] > 0, adm
    assert 0 < adm["peak_reserved_bytes"] <= cap and budget.used == 0
    _check_alone(rings, reqs)


def test_admission_rejects_what_never_fits(rings):
    """A budget smaller than one request's working memory for some operations:
    those requests are rejected (with the reason), the rest complete and equal
    their alone-runs."""
    from kunpbfv_amd import serving
    cap = max(rings[ln].batch_bytes("add", 2) for ln in rings)
    budget = serving.MemoryBudget(capacity=cap)
    reqs = serving.poisson_requests(qps=1e5, duration_s=4e-4, log_ns=[12, 13], seed=7, pool=6)
    stats = serving.serve(rings, reqs, keep=True, budget=budget)
    never = {(ln, op) for ln in rings for op in serving.OPS if rings[ln].batch_bytes(op, 1) > cap}
    assert never and any(op == "add" for _ln, op in {(r.log_n, r.op) for r in reqs})
    for r in reqs:
        assert bool(r.rejected) == ((r.log_n, r.op) in never), (r.rid, r.op, r.rejected)
    done = [r for r in reqs if not r.rejected]
    assert stats["admission"]["rejected"] == len(reqs) - len(done) > 0
    assert stats["requests"] == len(done) > 0
    _check_alone(rings, done)
