"""Continuous-batching request loop for mixed BFV work on one GPU.

Requests -- HomMult+relinearisation, encryption, RGSW x RLWE external product
and addition -- arrive at seeded Poisson times for several ring sizes.  The
loop keeps one queue per (ring, operation) and one HIP stream per ring; as
soon as a ring's stream has room it takes everything pending for one
operation (up to that ring's batch cap) and issues it as ONE batched C-ABI
call, so batches grow by themselves while the GPU is busy and rings of
different sizes run side by side.  Completion is timed with HIP events against
a start event, so a request's latency is (device completion - arrival).

Admission is by device memory.  Keys, RGSW ciphertexts, the operand pools
and each ring's workspace stay resident; what a batch needs (its workspace
share, the staged operands and its output) is reserved from one budget shared
by every ring -- by default the device memory still free once the rings are
resident, less a reserve -- and held until the batch's completion event
fires.  A batch that
does not fit is halved until it does; if not even one request fits, the
requests stay queued (deferred) until running batches release memory, and a
request that could not fit into the whole budget is rejected.

``serial_plain`` runs the same request list the way the reference's drivers
do (include/KeySwitch_23.cuh:244 per call): one request at a time, batch 1,
waiting for each to finish.

Ring sizes and chains follow the HE-standard 128-bit bounds for 60-bit primes:
N = 2^13: 2 + 1 primes (180 bits <= 218), 2^14: 6 + 1 (420 <= 438),
2^15: 13 + 1 (840 <= 881), 2^16: 28 + 1 (1740 bits); 4 auxiliary T primes,
gamma 2, t = 65537.
"""
import collections
import time

import numpy as np

import kunpbfv_amd as kb

OPS = ("mult", "encrypt", "extprod", "add")
DEFAULT_MIX = {"mult": 0.4, "encrypt": 0.3, "extprod": 0.2, "add": 0.1}
RING_CHAINS = {13: 2, 14: 6, 15: 13, 16: 28}
RGSW_BASE_BITS = 20


class Request:
    __slots__ = ("rid", "log_n", "op", "arrival", "a", "b", "done_ev", "latency_ms", "out", "rejected")

    def __init__(self, rid, log_n, op, arrival, a, b):
        self.rid, self.log_n, self.op, self.arrival, self.a, self.b = rid, log_n, op, arrival, a, b
        self.done_ev = None
        self.latency_ms = None
        self.out = None
        self.rejected = None  # reason, if admission refused the request


class MemoryBudget:
    """Bytes of device memory the loop may hold for in-flight batches.

    ``capacity`` defaults to the device memory free when the budget is made
    (after the rings' keys and operand pools are resident), less ``reserve``
    (a fraction kept back for the allocator's rounding and other users)."""

    def __init__(self, capacity=None, reserve=0.1):
        import torch
        if capacity is None:
            free, _total = torch.cuda.mem_get_info()
            capacity = int(free * (1.0 - reserve))
        self.capacity = int(capacity)
        self.used = 0
        self.peak = 0

    def fits(self, nbytes):
        return self.used + nbytes <= self.capacity

    def reserve(self, nbytes):
        if not self.fits(nbytes):
            return False
        self.used += nbytes
        self.peak = max(self.peak, self.used)
        return True

    def release(self, nbytes):
        self.used -= nbytes
        assert self.used >= 0


def poisson_requests(qps, duration_s, log_ns, seed, pool, mix=None):
    """Seeded Poisson arrivals; ring size uniform over log_ns, operation by mix;
    operands are indices into each ring's pool of prepared inputs."""
    mix = mix or DEFAULT_MIX
    rng = np.random.default_rng(seed)
    ops = list(mix)
    probs = np.array([mix[o] for o in ops], dtype=float)
    probs /= probs.sum()
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / qps)
        if t >= duration_s:
            return out
        out.append(Request(len(out), int(rng.choice(log_ns)), ops[rng.choice(len(ops), p=probs)], t,
                           int(rng.integers(pool)), int(rng.integers(pool))))


class Ring:
    """One ring size: context, keys, a pool of prepared operands (all resident
    in HBM), its stream and batch cap."""

    def __init__(self, log_n, qn=None, pool=16, max_batch=64, seed=1):
        import torch
        self.torch = torch
        qn = qn or RING_CHAINS[log_n]
        self.ctx = ctx = kb.Context.chain60(log_n, qn, 1, 4, 2, 65537)
        self.log_n = log_n
        s = bytes([seed, log_n]) + bytes(30)
        self.sk = ctx.keygen_sk(s)
        self.pk = ctx.keygen_pk(bytes([seed, log_n, 1]) + bytes(29), self.sk)
        self.rlk = ctx.keygen_rlk(bytes([seed, log_n, 2]) + bytes(29), self.sk)
        mg = np.zeros(ctx.n, dtype=np.uint64)
        mg[1] = 1  # RGSW(X): the external product is a negacyclic shift by one
        self.rgsw = ctx.rgsw_encrypt(bytes([seed, log_n, 3]) + bytes(29), self.sk, kb.to_device(mg[None])[0],
                                     RGSW_BASE_BITS)
        rng = np.random.default_rng(1000 + log_n)
        self.msgs = rng.integers(0, ctx.plain, size=(pool, ctx.n), dtype=np.uint64)
        md = kb.to_device(self.msgs)
        ctx.encode(md)
        self.pt = md  # encoded plaintexts [pool][N]
        self.cts = ctx.encrypt(bytes([seed, log_n, 4]) + bytes(29), self.pk, md)  # [pool][2][qn][N]
        self.ct_bytes = 2 * ctx.qn * ctx.n * 8
        self.max_batch = int(max_batch)
        self.stream = torch.cuda.Stream()
        self.enc_calls = 0
        # the ring's workspace: one grow-only buffer on its stream (calls on
        # one stream run in order, so they share it), sized here for the
        # largest batch so that no allocation happens while requests wait
        self._ws = None
        self._stream_ws(max(self.workspace_bytes(op, self.max_batch) for op in OPS))

    def _stream_ws(self, nbytes):
        torch = self.torch
        if self._ws is None or self._ws.numel() < nbytes:
            with torch.cuda.stream(self.stream):
                self._ws = None
                self._ws = torch.empty(max(int(nbytes), 256), dtype=torch.uint8, device="cuda")
        return self._ws

    def workspace_bytes(self, op, batch):
        lib, h = kb.lib, self.ctx.h
        if op == "mult":
            return self.ctx.mult_workspace_bytes(batch)
        if op == "extprod":
            return lib.kbfv_external_product_workspace_bytes(h, batch, RGSW_BASE_BITS)
        if op == "encrypt":
            return lib.kbfv_encrypt_workspace_bytes(h, batch)
        return 0

    def batch_bytes(self, op, batch):
        """Device memory one call of `batch` requests holds beyond the resident
        data: workspace, staged operands and output (each rounded up to the
        allocator's 512-byte granule)."""
        def r(x):
            return -(-x // 512) * 512
        staged = {"mult": 2, "add": 2, "extprod": 1, "encrypt": 0}[op] * batch * self.ct_bytes
        if op == "encrypt":
            staged = batch * self.ctx.n * 8  # the plaintexts
        return r(self.workspace_bytes(op, batch)) + r(staged) + r(batch * self.ct_bytes) + 4096

    def run(self, op, reqs, stream=None, keep=False, hold=None):
        """One batched call for the requests (same op), on `stream` (default:
        the ring's, whose workspace it shares with the ring's other calls).
        Returns the output batch; `hold`, if given, receives every tensor the
        call uses so the caller can keep them alive until the call completes."""
        torch, ctx = self.torch, self.ctx
        st = stream or self.stream
        wsb = self.workspace_bytes(op, len(reqs))
        ws = self._stream_ws(wsb) if st is self.stream else None
        with torch.cuda.stream(st):
            if ws is None:  # a call on another stream gets its own workspace
                ws = torch.empty(max(wsb, 256), dtype=torch.uint8, device="cuda")
            ia = torch.tensor([r.a for r in reqs], dtype=torch.long, device="cuda")
            if op == "encrypt":
                m = self.pt.index_select(0, ia)
                out = ctx.empty(len(reqs), 2, ctx.qn, ctx.n)
                self.enc_calls += 1
                seed = bytes([9, self.log_n]) + self.enc_calls.to_bytes(8, "little") + bytes(22)
                kb.check(kb.lib.kbfv_encrypt(ctx.h, kb._seed(seed), kb.ptr(self.pk), kb.ptr(m), kb.ptr(out),
                                             len(reqs), kb.ptr(ws), ws.numel(), kb.stream_ptr(st)))
                staged = (m,)
            else:
                c1 = self.cts.index_select(0, ia)
                out = ctx.empty(*c1.shape)
                staged = (c1,)
                if op == "extprod":
                    kb.check(kb.lib.kbfv_external_product(ctx.h, kb.ptr(self.rgsw), kb.ptr(c1), kb.ptr(out),
                                                          len(reqs), RGSW_BASE_BITS, kb.ptr(ws),
                                                          ws.numel(), kb.stream_ptr(st)))
                else:
                    ib = torch.tensor([r.b for r in reqs], dtype=torch.long, device="cuda")
                    c2 = self.cts.index_select(0, ib)
                    staged = (c1, c2)
                    if op == "mult":
                        kb.check(kb.lib.kbfv_mult_relin(ctx.h, kb.ptr(self.rlk), kb.ptr(c1), kb.ptr(c2),
                                                        kb.ptr(out), len(reqs), kb.ptr(ws), ws.numel(),
                                                        kb.stream_ptr(st)))
                    else:
                        kb.check(kb.lib.kbfv_add(ctx.h, kb.ptr(c1), kb.ptr(c2), kb.ptr(out), len(reqs),
                                                 kb.stream_ptr(st)))
            ev = torch.cuda.Event(enable_timing=True)
            ev.record(st)
        if hold is not None:
            hold.extend((ws, ia, out) + staged)
        for i, r in enumerate(reqs):
            r.done_ev = ev
            if keep:
                r.out = out[i]
        return out


def _finish(reqs, start_ev, t_host0_to_dev_ms=0.0):
    lat = []
    for r in reqs:
        if r.rejected:
            continue
        r.latency_ms = start_ev.elapsed_time(r.done_ev) - r.arrival * 1e3 + t_host0_to_dev_ms
        lat.append(r.latency_ms)
    lat = np.sort(np.array(lat if lat else [0.0]))
    return lat


def _summary(reqs, lat, wall_s):
    pct = lambda p: float(lat[min(len(lat) - 1, int(round(p / 100 * (len(lat) - 1))))])  # noqa: E731
    done = sum(1 for r in reqs if not r.rejected)
    return {"requests": done, "throughput_rps": round(done / wall_s, 2),
            "p50_ms": round(pct(50), 3), "p99_ms": round(pct(99), 3), "wall_s": round(wall_s, 3)}


def serve(rings, reqs, max_inflight=2, keep=False, budget=None):
    """Continuous batching over the request list (sorted by arrival), admitted
    against `budget` (a MemoryBudget; default: the device memory free now,
    less 10%)."""
    import torch
    torch.cuda.synchronize()
    budget = budget if budget is not None else MemoryBudget()
    start_ev = torch.cuda.Event(enable_timing=True)
    start_ev.record()
    t0 = time.perf_counter()
    queues = collections.defaultdict(collections.deque)
    inflight = collections.defaultdict(collections.deque)  # log_n -> (event, reserved bytes) per batch
    batches = collections.Counter()
    deferrals = halvings = rejected = 0
    nxt = 0

    def retire(fl):
        while fl and fl[0][0].query():
            budget.release(fl.popleft()[1])

    while nxt < len(reqs) or any(queues.values()):
        now = time.perf_counter() - t0
        while nxt < len(reqs) and reqs[nxt].arrival <= now:
            r = reqs[nxt]
            queues[(r.log_n, r.op)].append(r)
            nxt += 1
        issued = False
        for (log_n, op), q in queues.items():
            if not q:
                continue
            fl = inflight[log_n]
            retire(fl)
            if len(fl) >= max_inflight:
                continue
            ring = rings[log_n]
            b = min(len(q), ring.max_batch)
            need = ring.batch_bytes(op, b)
            while b > 1 and not budget.fits(need):
                b = (b + 1) // 2
                need = ring.batch_bytes(op, b)
                halvings += 1
            if not budget.reserve(need):
                if need > budget.capacity:  # cannot fit even with nothing else in flight
                    r = q.popleft()
                    r.rejected = f"needs {need} bytes, budget {budget.capacity}"
                    rejected += 1
                else:
                    for other in inflight.values():  # memory may have come free on another ring
                        retire(other)
                    deferrals += 1
                continue
            take = [q.popleft() for _ in range(b)]
            # the call's tensors live on the ring's stream: the allocator
            # reuses their memory only for later work on that stream, so the
            # reservation (not the tensors) is what is held until completion
            ring.run(op, take, keep=keep)
            fl.append((take[0].done_ev, need))
            batches[(log_n, op)] += 1
            issued = True
        if not issued:
            wait = (reqs[nxt].arrival - (time.perf_counter() - t0)) if nxt < len(reqs) else 0.0
            time.sleep(min(max(wait, 0.0), 2e-4) if wait > 0 else 5e-5)
    torch.cuda.synchronize()
    for fl in inflight.values():
        retire(fl)
    wall = time.perf_counter() - t0
    lat = _finish(reqs, start_ev)
    out = _summary(reqs, lat, wall)
    out["batches"] = int(sum(batches.values()))
    out["mean_batch"] = round(out["requests"] / max(1, out["batches"]), 2)
    out["admission"] = {"budget_bytes": budget.capacity, "peak_reserved_bytes": budget.peak,
                        "deferrals": deferrals, "halvings": halvings, "rejected": rejected}
    return out


def serial_plain(rings, reqs, keep=False):
    """The same requests one at a time, batch 1, each waited for."""
    import torch
    torch.cuda.synchronize()
    start_ev = torch.cuda.Event(enable_timing=True)
    start_ev.record()
    t0 = time.perf_counter()
    st = torch.cuda.current_stream()
    for r in reqs:
        wait = r.arrival - (time.perf_counter() - t0)
        if wait > 0:
            time.sleep(wait)
        rings[r.log_n].run(r.op, [r], stream=st, keep=keep)
        r.done_ev.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    lat = _finish(reqs, start_ev)
    return _summary(reqs, lat, wall)


def copy_requests(reqs):
    return [Request(r.rid, r.log_n, r.op, r.arrival, r.a, r.b) for r in reqs]
