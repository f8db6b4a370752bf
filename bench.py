#!/usr/bin/env python3
"""Benchmark: batched BFV HomMult + relinearisation at N=2^16 on one MI355X.

Workload (one step): 128 HomMult+relin of ciphertext pairs, N = 2^16, 28 x
60-bit ciphertext primes + 1 special prime, t = 65537 (coefficient encoding:
65537 is not 1 mod 2^17), relinearisation key with dnum = 28 digits held in a
4-prime auxiliary basis T (gamma = 2).  Inputs are fresh encryptions of seeded
uniform plaintexts, resident in HBM before the timed region.

Prints one JSON line: throughput of the batched path (HomMult+relin/s), the
plain per-ciphertext path on the same inputs (128 launches of batch 1), their
ratio, step-latency percentiles, a roofline entry for the dominant kernel
(timed with HIP events on the library's stream) and the CPU oracle's time for
one HomMult+relin at the same shape on this host.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E

# Measured HBM traffic of one inverse NTT over the QR basis of the whole batch
# (the roofline entry), keyed by (log N, Q primes, batch, plan word): rocprofv3
# FETCH_SIZE and WRITE_SIZE runs over `bench.py --roofline-only`, summed over
# the transform's launches and scaled by the factor each counter reads for a
# 4 GiB device copy in the same run.  A plan without a recorded profile
# reports traffic null rather than a figure measured for another plan.
NTT_TRAFFIC = {
    (16, 28, 128, 0x3cd8): (2 * 23_000_583_594, "profiles/ntt_qr_inverse_n65536_b128_counters.md"),
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=128)
    ap.add_argument("--log-n", type=int, default=16)
    ap.add_argument("--qn", type=int, default=28)
    ap.add_argument("--plain-steps", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-graph", action="store_true", help="skip the N=2^15 HIP-graph measurement")
    ap.add_argument("--no-serving", action="store_true", help="skip the mixed-N serving measurement")
    ap.add_argument("--serving-qps", type=float, default=3000.0)
    ap.add_argument("--serving-seconds", type=float, default=3.0)
    ap.add_argument("--no-autotune", action="store_true", help="keep the built-in NTT plans")
    ap.add_argument("--no-ntt-alone", action="store_true",
                    help="skip the NTT-alone measurement at N=2^14, 8 limbs, 256 polynomials")
    ap.add_argument("--no-mixed", action="store_true",
                    help="skip the HomMult + RGSW external-product mix at the same shape")
    ap.add_argument("--roofline-only", action="store_true",
                    help="only the roofline kernel and a 4 GiB copy (for counter runs)")
    return ap.parse_args()


def pct(xs, p):
    xs = sorted(xs)
    k = min(len(xs) - 1, max(0, int(round(p / 100.0 * (len(xs) - 1)))))
    return xs[k]


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(ctx, ct1, ct2, rlk_shape):
    """Oracle HomMult+relin on one ciphertext pair of the same shape: once on one
    core, then one independent product per host thread (ctypes releases the
    interpreter lock, the oracle keeps no global state)."""
    import numpy as np
    from concurrent.futures import ThreadPoolExecutor
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as orc
    import kunpbfv_amd as kb
    oc = orc.Ctx(ctx.log_n, ctx._q, ctx._p, ctx._t, ctx.gamma, ctx.plain)
    rng = np.random.default_rng(5)
    # timing does not depend on the key's values: a random key of the right shape
    rlk = np.empty(rlk_shape, dtype=np.uint64)
    tm = oc.mod[oc.oT:oc.oT + oc.tn]
    for k in range(oc.tn):
        rlk[..., k, :] = rng.integers(0, int(tm[k]), size=rlk[..., k, :].shape, dtype=np.uint64)
    a, b = kb.to_host(ct1[0]), kb.to_host(ct2[0])
    t0 = time.perf_counter()
    oc.mult_relin(rlk, a, b)
    dt1 = time.perf_counter() - t0
    threads = host_threads()
    t0 = time.perf_counter()
    with ThreadPoolExecutor(threads) as ex:
        list(ex.map(lambda _: oc.mult_relin(rlk, a, b), range(threads)))
    dtn = time.perf_counter() - t0
    return {"value": round(threads / dtn, 4), "unit": "HomMult+relin/s", "cores": threads, "kind": "port",
            "sample": f"{threads} independent HomMult+relin at N=2^{ctx.log_n}, {ctx.qn}+{ctx.pn} primes, "
                      f"CPU oracle, one per thread, {dtn:.2f} s; one core alone: {dt1:.2f} s per op",
            "single_core_value": round(1.0 / dt1, 4), "cpu_model": cpu_model(),
            "nproc": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0))}


def host_threads():
    """Threads for the CPU baseline: every CPU this process may run on, capped
    at the 16 host cores a one-GPU box allots to its GPU's jobs."""
    return max(1, min(len(os.sched_getaffinity(0)), 16))


def cpu_c1(threads, seconds=1.0):
    """The CPU-runnable BFV workload at N=4096 (two 36-bit Q primes and one
    37-bit special prime, 1 mod 8192, t=65537, batching): encrypt, add,
    mult+relin and decrypt of seeded uniform slot vectors, each timed on the
    CPU oracle for about `seconds` on one core and then with one independent
    stream of operations per host thread.  The reference's own timing loops
    (perf_encrypt.cu:26-40, perf_add.cu:31-51, perf_mult.cu:29-45) time the
    same four operations one call at a time."""
    import numpy as np
    from concurrent.futures import ThreadPoolExecutor
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as orc
    import params as tp
    d = tp.c1_params()
    oc = orc.Ctx(d["log_n"], d["q"], d["p"], d["t"], d["gamma"], d["plain"])
    rng = np.random.default_rng(4096)
    sk = oc.keygen_sk(bytes([71]) * 32)
    pk = oc.keygen_pk(bytes([72]) * 32, sk)
    rlk = oc.keygen_rlk(bytes([73]) * 32, sk)
    m1, m2 = (oc.encode(rng.integers(0, d["plain"], size=oc.n, dtype=np.uint64)) for _ in range(2))
    c1 = oc.encrypt(bytes([74]) * 32, pk, m1)
    c2 = oc.encrypt(bytes([75]) * 32, pk, m2, entry=1)
    qv = oc.mod[oc.oQ:oc.oQ + oc.qn].reshape(1, oc.qn, 1)
    ops = {
        "encrypt": lambda: oc.encrypt(bytes([76]) * 32, pk, m1),
        "add": lambda: (c1 + c2) % qv,
        "mult_relin": lambda: oc.mult_relin(rlk, c1, c2),
        "decrypt": lambda: oc.decrypt(sk, c1),
    }
    # the product must decrypt to the slot-wise product
    dec = oc.decode(oc.decrypt(sk, oc.mult_relin(rlk, c1, c2)))
    s1, s2 = oc.decode(m1), oc.decode(m2)
    ok = bool(np.array_equal(dec, (s1.astype(object) * s2.astype(object) % d["plain"]).astype(np.uint64)))

    def rate(fn, secs):
        fn()
        k, t0 = 0, time.perf_counter()
        while True:
            fn()
            k += 1
            dt = time.perf_counter() - t0
            if dt >= secs:
                return k / dt

    single = {name: round(rate(fn, seconds), 2) for name, fn in ops.items()}
    multi = {}
    with ThreadPoolExecutor(threads) as ex:
        for name, fn in ops.items():
            if name == "add":  # numpy element-wise: holds the interpreter lock, not threaded
                continue
            multi[name] = round(sum(ex.map(lambda _: rate(fn, seconds), range(threads))), 2)
    return {"workload": "BFV at N=4096, Q = 2 x 36-bit primes, P = 1 x 37-bit prime, t=65537 (batching), "
                        "seeded uniform slots, ternary secret, CBD error (sigma 3.24)",
            "unit": "op/s", "kind": "port", "single_core": single, "threads": threads,
            "all_threads": multi, "mult_decrypts_to_slot_product": ok,
            "sample": f"each op repeated for {seconds:.1f} s on one core, then on {threads} threads"}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU; multi-GPU runs are not part of "
                                   "this benchmark (the batch shards by ciphertext across GPUs)"}))
        sys.exit(2)
    import numpy as np
    import torch
    import kunpbfv_amd as kb

    torch.cuda.init()
    ctx = kb.Context.chain60(args.log_n, args.qn, 1, 4, 2, 65537)
    B, N = args.batch, ctx.n
    stream = torch.cuda.current_stream()
    tuned = None
    if not args.no_autotune:
        tuned = autotune_report(ctx, ctx.autotune(B))
        torch.cuda.empty_cache()
    if args.roofline_only:
        # calibration: a plain device copy of 4 GiB (4 GiB read + 4 GiB written)
        src = torch.empty(1 << 29, dtype=torch.int64, device="cuda")
        dst = torch.empty_like(src)
        src.fill_(1)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
        print(json.dumps({"roofline": kernel_roofline(ctx, B, stream), "copy_bytes": 1 << 32}), flush=True)
        return
    sk = ctx.keygen_sk(bytes([1]) * 32)
    pk = ctx.keygen_pk(bytes([2]) * 32, sk)
    rlk = ctx.keygen_rlk(bytes([3]) * 32, sk)
    rng = np.random.default_rng(2024)
    m1 = rng.integers(0, ctx.plain, size=(B, N), dtype=np.uint64)
    m2 = rng.integers(0, ctx.plain, size=(B, N), dtype=np.uint64)
    # entries 0 and 1 of the second batch are a monomial and a constant so that
    # the decrypted products can be checked exactly
    m2[0] = 0
    m2[0, 5] = 1
    m2[1] = 0
    m2[1, 0] = 3
    ct1 = ctx.encrypt(bytes([4]) * 32, pk, kb.to_device(m1))
    ct2 = ctx.encrypt(bytes([5]) * 32, pk, kb.to_device(m2))
    out = ctx.empty(*ct1.shape)
    ws_bytes = ctx.mult_workspace_bytes(B)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device="cuda")

    # ---- correctness: decrypt the two checkable products
    ctx.mult_relin(rlk, ct1, ct2, out=out, ws=ws)
    dec = kb.to_host(ctx.decrypt(sk, out[0:2].contiguous()))
    want0 = np.zeros(N, dtype=np.int64)
    want0[5:] = m1[0, :-5].astype(np.int64)
    want0[:5] = -m1[0, -5:].astype(np.int64)
    ok = np.array_equal(dec[0].astype(np.int64) % ctx.plain, want0 % ctx.plain)
    ok &= np.array_equal(dec[1], (m1[1] * 3) % ctx.plain)
    if not ok:
        print(json.dumps({"error": "decrypted HomMult result is wrong"}))
        sys.exit(1)
    ref_out = out.clone()

    # ---- batched path
    for _ in range(args.warmup):
        ctx.mult_relin(rlk, ct1, ct2, out=out, ws=ws)
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
    t0 = time.perf_counter()
    ev[0].record(stream)
    for i in range(args.steps):
        ctx.mult_relin(rlk, ct1, ct2, out=out, ws=ws)
        ev[i + 1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    step_ms = [ev[i].elapsed_time(ev[i + 1]) for i in range(args.steps)]
    ms_per_step = sum(step_ms) / len(step_ms)
    batched_ops = B / (ms_per_step / 1e3)
    bit_exact = bool(torch.equal(out, ref_out))

    # ---- plain per-ciphertext path (128 launches of batch 1) on the same inputs
    plain_ops = None
    plain_exact = None
    if not args.no_plain:
        out1 = ctx.empty(*ct1.shape)
        ws1 = torch.empty(ctx.mult_workspace_bytes(1), dtype=torch.uint8, device="cuda")
        c1s = [ct1[i:i + 1] for i in range(B)]
        c2s = [ct2[i:i + 1] for i in range(B)]
        o1s = [out1[i:i + 1] for i in range(B)]
        for i in range(B):  # warm-up pass
            ctx.mult_relin(rlk, c1s[i], c2s[i], out=o1s[i], ws=ws1)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(args.plain_steps):
            for i in range(B):
                ctx.mult_relin(rlk, c1s[i], c2s[i], out=o1s[i], ws=ws1)
        e1.record(stream)
        torch.cuda.synchronize()
        plain_ms = e0.elapsed_time(e1) / args.plain_steps
        plain_ops = B / (plain_ms / 1e3)
        plain_exact = bool(torch.equal(out1, ref_out))

    # ---- dominant kernel: the NTT pass, measured on the pipeline's largest launch
    # (inverse transform over the QR basis of the 3 tensor polynomials of every ciphertext)
    roof = kernel_roofline(ctx, B, stream)

    result = {
        "metric": "HomMult+relin/s at N=2^16 batched vs plain (bit-exact); p50/p99 latency",
        "value": round(batched_ops, 2),
        "unit": "HomMult+relin/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic (seeded uniform plaintexts mod t, fresh encryptions)",
        "config": {"workload": f"batched HomMult+relin, N=2^{args.log_n}, {args.qn}+1 x 60-bit primes, "
                               f"t=65537, dnum={ctx.dnum}, T basis {ctx.tn} primes (gamma={ctx.gamma}), "
                               f"batch {B}",
                   "batch": B, "log_n": args.log_n, "q_primes": args.qn, "special_primes": 1},
        "plain_path_ops_per_s": round(plain_ops, 2) if plain_ops else None,
        "batched_over_plain": round(batched_ops / plain_ops, 3) if plain_ops else None,
        "bit_exact_batched_vs_plain": plain_exact,
        "bit_exact_across_steps": bit_exact,
        "step_latency_ms": {"p50": round(pct(step_ms, 50), 3), "p99": round(pct(step_ms, 99), 3)},
        "wall_s": round(wall, 3),
        "roofline": roof,
        "ntt_autotune": tuned,
    }
    if not args.no_mixed:
        result["mixed_ext_product"] = mixed_c4(ctx, sk, rlk, ct1, ct2, out, ws, m1, stream)
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(ctx, ct1, ct2, tuple(rlk.shape))
        result["cpu_baseline"]["c1_n4096"] = cpu_c1(host_threads())
    del ws, out, ct1, ct2, rlk
    if not args.no_ntt_alone:
        result["ntt_n16384_l8_b256"] = ntt_alone(stream)
    if not args.no_graph:
        result["graph_n32768_b64"] = graph_c3(stream)
    if not args.no_serving:
        result["serving"] = serving_mix(args.serving_qps, args.serving_seconds)
    print(json.dumps(result), flush=True)


def mixed_c4(ctx, sk, rlk, ct1, ct2, out, ws, m1, stream, bits=20):
    """The batch's HomMult+relin mixed with as many RGSW x RLWE external
    products (gadget base 2^bits over each Q limb, one RGSW of X^3 shared by
    the batch, applied to the first input batch): each alone and both in one
    timed region; the first product must decrypt to m1[0] * X^3."""
    import numpy as np
    import torch
    import kunpbfv_amd as kb
    B, N = ct1.shape[0], ctx.n
    mg = np.zeros(N, dtype=np.uint64)
    mg[3] = 1
    rgsw = ctx.rgsw_encrypt(bytes([6]) * 32, sk, kb.to_device(mg), bits)
    xout = ctx.empty(*ct1.shape)
    xws = torch.empty(kb.lib.kbfv_external_product_workspace_bytes(ctx.h, B, bits), dtype=torch.uint8,
                      device="cuda")

    def ext():
        kb.check(kb.lib.kbfv_external_product(ctx.h, kb.ptr(rgsw), kb.ptr(ct1), kb.ptr(xout), B, bits,
                                              kb.ptr(xws), xws.numel(), kb.stream_ptr(stream)))

    def timed(fn):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1)

    ext()
    torch.cuda.synchronize()
    dec = kb.to_host(ctx.decrypt(sk, xout[0:1].contiguous()))[0]
    want = np.zeros(N, dtype=np.int64)
    want[3:] = m1[0, :-3].astype(np.int64)
    want[:3] = -m1[0, -3:].astype(np.int64)
    ok = bool(np.array_equal(dec.astype(np.int64) % ctx.plain, want % ctx.plain))
    t_ext = timed(ext)
    t_mult = timed(lambda: ctx.mult_relin(rlk, ct1, ct2, out=out, ws=ws))

    def both():
        ctx.mult_relin(rlk, ct1, ct2, out=out, ws=ws)
        ext()

    t_mix = timed(both)
    del rgsw, xws, xout
    torch.cuda.empty_cache()
    return {"workload": f"{B} HomMult+relin + {B} RGSW x RLWE external products (gadget base 2^{bits}, "
                        f"{ctx.qn} x {60 // bits} digits per component), N=2^{ctx.log_n}",
            "ext_product_ms": round(t_ext, 3), "ext_products_per_s": round(B / (t_ext / 1e3), 2),
            "hommult_ms": round(t_mult, 3), "mixed_ms": round(t_mix, 3),
            "mixed_ops_per_s": round(2 * B / (t_mix / 1e3), 2), "ext_product_decrypts_to_shift": ok}


def ntt_alone(stream, polys=256, reps=10):
    """The batched NTT kernel alone: N=2^14, 8 x 60-bit primes (1 mod 2^17),
    256 polynomials of seeded uniform residues, forward then inverse with the
    autotuned plans; algorithmic bytes = every limb read and written once per
    transform.  The inverse must return the input bit for bit."""
    import numpy as np
    import torch
    import kunpbfv_amd as kb
    ctx = kb.Context.chain60(14, 8, 1, 4, 2, 65537)
    mod0, limbs = ctx.oQ, ctx.qn
    rng = np.random.default_rng(2)
    x = np.empty((polys, limbs, ctx.n), dtype=np.uint64)
    for i in range(limbs):
        x[:, i, :] = rng.integers(0, int(ctx.mod[mod0 + i]), size=(polys, ctx.n), dtype=np.uint64)
    xd = kb.to_device(x)
    plans = {d: ctx.autotune_ntt(d, mod0, limbs, polys, reps=3)[0] for d in (+1, -1)}
    yd = xd.clone()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    ctx.ntt(yd, yd, +1, mod0, limbs, polys=polys)
    ctx.ntt(yd, yd, -1, mod0, limbs, polys=polys)
    torch.cuda.synchronize()
    ev[0].record(stream)
    for _ in range(reps):
        ctx.ntt(yd, yd, +1, mod0, limbs, polys=polys)
    ev[1].record(stream)
    for _ in range(reps):
        ctx.ntt(yd, yd, -1, mod0, limbs, polys=polys)
    ev[2].record(stream)
    torch.cuda.synchronize()
    t_f = ev[0].elapsed_time(ev[1]) / reps
    t_i = ev[1].elapsed_time(ev[2]) / reps
    exact = bool(torch.equal(yd, xd))
    nbytes = 2 * polys * limbs * ctx.n * 8  # read + write of every limb, per transform
    # the single-pass plan (one workgroup per limb) for comparison
    bucket = kb.lib.kbfv_ntt_plan_bucket(limbs * polys)
    ctx.set_ntt_plan(+1, bucket, ctx.log_n | 1 << 12)
    ev[0].record(stream)
    for _ in range(reps):
        ctx.ntt(yd, yd, +1, mod0, limbs, polys=polys)
    ev[1].record(stream)
    torch.cuda.synchronize()
    t_single = ev[0].elapsed_time(ev[1]) / reps
    ctx.set_ntt_plan(+1, bucket, plans[+1])

    def fields(p):
        cs, ct, rt, cx, rx = kb.ntt_plan_fields(p, ctx.log_n)
        return {"plan": hex(p), "passes": 1 if cs == ctx.log_n else 2, "split_log": [cs, ctx.log_n - cs],
                "tile_log": [ct, rt]}

    res = {"workload": f"negacyclic NTT, N=2^14, {limbs} x 60-bit primes, {polys} polynomials "
                       f"({polys * limbs} limbs), forward then inverse, {reps} of each",
           "forward_ms": round(t_f, 4), "inverse_ms": round(t_i, 4),
           "forward_gbs": round(nbytes / t_f / 1e6, 1), "inverse_gbs": round(nbytes / t_i / 1e6, 1),
           "forward_frac": round(nbytes / t_f / 1e6 / HBM_PEAK_GBS, 4),
           "inverse_frac": round(nbytes / t_i / 1e6 / HBM_PEAK_GBS, 4),
           "per_launch_gbs": {"forward": round(nbytes * fields(plans[+1])["passes"] / t_f / 1e6, 1),
                              "inverse": round(nbytes * fields(plans[-1])["passes"] / t_i / 1e6, 1)},
           "single_pass_forward_ms": round(t_single, 4),
           "bytes_per_transform": nbytes, "forward_plan": fields(plans[+1]), "inverse_plan": fields(plans[-1]),
           "working_set_bytes": polys * limbs * ctx.n * 8,
           "inverse_of_forward_is_identity": exact}
    del xd, yd
    torch.cuda.empty_cache()
    return res


def graph_c3(stream, batch=64, reps=5):
    """N=2^15, 13 + 1 x 60-bit primes, t=65537 (batching), dnum=13, batch 64:
    the batched pipeline captured as one HIP graph, replayed, against the same
    batch as one eager batched call and as 64 launches of batch 1."""
    import numpy as np
    import torch
    import kunpbfv_amd as kb
    ctx = kb.Context.chain60(15, 13, 1, 4, 2, 65537)
    ctx.autotune(batch)
    sk = ctx.keygen_sk(bytes([7]) * 32)
    pk = ctx.keygen_pk(bytes([8]) * 32, sk)
    rlk = ctx.keygen_rlk(bytes([9]) * 32, sk)
    rng = np.random.default_rng(15)
    ms = []
    cts = []
    for seed in (10, 11):
        m = rng.integers(0, ctx.plain, size=(batch, ctx.n), dtype=np.uint64)
        md = kb.to_device(m)
        ctx.encode(md)
        ms.append(m)
        cts.append(ctx.encrypt(bytes([seed]) * 32, pk, md))
    c1, c2 = cts
    nb = ctx.mult_workspace_bytes(batch)
    ws = torch.empty(nb, dtype=torch.uint8, device="cuda")
    out_e, out_g, out_p = (ctx.empty(*c1.shape) for _ in range(3))
    graph = ctx.mult_relin_graph(rlk, c1, c2, out_g, ws)
    ws1 = torch.empty(ctx.mult_workspace_bytes(1), dtype=torch.uint8, device="cuda")

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    def plain():
        for i in range(batch):
            ctx.mult_relin(rlk, c1[i:i + 1], c2[i:i + 1], out=out_p[i:i + 1], ws=ws1)

    t_e = timed(lambda: ctx.mult_relin(rlk, c1, c2, out=out_e, ws=ws))
    t_g = timed(lambda: graph.launch(stream))
    t_p = timed(plain)
    dec = ctx.decrypt(sk, out_g)
    ctx.decode(dec)
    want = (ms[0].astype(object) * ms[1].astype(object)) % ctx.plain
    ok = np.array_equal(kb.to_host(dec).astype(object), want)
    res = {"workload": "HomMult+relin, N=2^15, 13+1 x 60-bit primes, t=65537, dnum=13, batch 64",
           "graph_ops_per_s": round(batch / (t_g / 1e3), 2),
           "eager_batched_ops_per_s": round(batch / (t_e / 1e3), 2),
           "plain_ops_per_s": round(batch / (t_p / 1e3), 2),
           "graph_over_plain": round(t_p / t_g, 3),
           "graph_ms": round(t_g, 3), "eager_ms": round(t_e, 3), "plain_ms": round(t_p, 3),
           "bit_exact_graph_eager_plain": bool(torch.equal(out_g, out_e) and torch.equal(out_g, out_p)),
           "decrypts_to_product": bool(ok)}
    del graph
    return res


def serving_mix(qps, seconds, seed=11):
    """Seeded Poisson arrivals over N = 2^13..2^16 (40% HomMult+relin, 30%
    encryption, 20% RGSW external product, 10% addition): the continuous-
    batching loop against the serial batch-1 loop on the same requests."""
    from kunpbfv_amd import serving
    rings = {ln: serving.Ring(ln, pool=16, max_batch=64) for ln in (13, 14, 15, 16)}
    reqs = serving.poisson_requests(qps, seconds, sorted(rings), seed, pool=16)
    batched = serving.serve(rings, reqs)
    plain = serving.serial_plain(rings, serving.copy_requests(reqs))
    return {"workload": f"Poisson {qps:g} req/s for {seconds:g} s, N=2^13..2^16 (2/6/13/28 + 1 x 60-bit), "
                        "40% mult+relin / 30% encrypt / 20% RGSW ext. product / 10% add, seed "
                        f"{seed}", "batched": batched, "serial_plain": plain}


def autotune_report(ctx, tuned):
    """The NTT plans the autotuner picked, per direction and transform-count bucket."""
    import kunpbfv_amd as kb
    rows = []
    for (direction, bucket), (plan, ms, transforms) in sorted(tuned.items()):
        cs, ct, rt, cx, rx = kb.ntt_plan_fields(plan, ctx.log_n)
        rows.append({"dir": "forward" if direction > 0 else "inverse", "bucket": bucket,
                     "transforms": transforms, "plan": hex(plan), "split_log": [cs, ctx.log_n - cs],
                     "tile_log": [ct, rt], "xcd_order": [cx, rx], "ms": round(ms, 4)})
    return rows


def kernel_roofline(ctx, B, stream):
    """The dominant kernel, k_ntt_pass, as the pipeline's largest NTT: one
    inverse negacyclic NTT over the QR basis of the 3 tensor polynomials of
    every ciphertext, timed with HIP events on the stream the library launches
    on.  Algorithmic bytes per transform = every limb read once and written
    once (2 x 512 KiB per 2^16-coefficient limb); a transform is two kernel
    launches (column and row pass) per chunk of its plan."""
    import torch
    import kunpbfv_amd as kb
    sn = ctx.qn + ctx.rn
    polys = 3 * B
    x = ctx.empty(polys, sn, ctx.n)
    x.random_(0, 1 << 40)
    reps = 5
    evs = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ctx.ntt(x, x, -1, ctx.oQ, sn, polys=polys)
    torch.cuda.synchronize()
    evs[0].record(stream)
    for _ in range(reps):
        ctx.ntt(x, x, -1, ctx.oQ, sn, polys=polys)
    evs[1].record(stream)
    torch.cuda.synchronize()
    per_transform_s = evs[0].elapsed_time(evs[1]) / reps / 1e3
    limbs = polys * sn
    bytes_per_transform = limbs * ctx.n * 8 * 2
    plan = ctx.ntt_plan(-1, kb.lib.kbfv_ntt_plan_bucket(limbs))
    chunk = kb.ntt_plan_chunk_bytes(plan)
    chunks = 1 if not chunk else -(-polys // max(1, chunk // (sn * ctx.n * 8)))
    launches = 2 * chunks
    achieved = bytes_per_transform / per_transform_s / 1e9
    traffic, source = NTT_TRAFFIC.get((ctx.log_n, ctx.qn, B, plan), (None, None))
    return {"kernel": "k_ntt_pass (inverse NTT over the QR basis, whole batch: column + row pass)",
            "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "traffic_source": source,
            "unit_of_work": f"one transform of {limbs} limbs of 2^{ctx.log_n} coefficients",
            "bytes_per_transform": bytes_per_transform, "transform_ms": round(per_transform_s * 1e3, 3),
            "launches_per_transform": launches, "avg_launch_ms": round(per_transform_s * 1e3 / launches, 4),
            "per_pass_frac": round(achieved * 2 / HBM_PEAK_GBS, 4),
            "ntt_plan": hex(plan), "chunk_bytes": chunk}


if __name__ == "__main__":
    main()
