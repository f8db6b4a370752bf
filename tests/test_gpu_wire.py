"""GPU: wire-format round trips through the C ABI for every object kind
(ciphertext batch, plaintexts, secret, public, relinearisation and Galois keys,
RGSW) — the loaded words equal the saved ones bit for bit and the payload is
the device words themselves — and the loader's rejections: wrong kind or
count, other parameters, a corrupted payload (checksum), a residue at or above
its modulus (with a consistent checksum), and a truncated blob."""
import struct

import numpy as np
import pytest

import params as P

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup():
    import kunpbfv_amd as kb
    d = P.chain60(12, 3, 1, 3, 1)
    ctx = kb.Context(d["log_n"], d["q"], d["p"], d["t"], d["gamma"], d["plain"])
    sk = ctx.keygen_sk(bytes([1]) * 32)
    pk = ctx.keygen_pk(bytes([2]) * 32, sk)
    rng = np.random.default_rng(4)
    m = rng.integers(0, ctx.plain, size=(5, ctx.n), dtype=np.uint64)
    md = kb.to_device(m)
    ctx.encode(md)
    ct = ctx.encrypt(bytes([3]) * 32, pk, md)
    return kb, ctx, sk, pk, m, ct


def roundtrip(kb, ctx, kind, objs, aux=0):
    blob = ctx.wire_save(kind, objs, aux)
    assert len(blob) == 64 + objs.numel() * 8
    assert kb.wire_peek(blob) == (kind, objs.shape[0], aux)
    payload = np.frombuffer(blob[64:], dtype=np.uint64)
    assert np.array_equal(payload, kb.to_host(objs).reshape(-1))
    back = ctx.empty(*objs.shape)
    back.fill_(-1)
    ctx.wire_load(blob, kind, back)
    assert np.array_equal(kb.to_host(back), kb.to_host(objs))
    return blob, back


def test_every_kind_round_trips(setup):
    kb, ctx, sk, pk, m, ct = setup
    _, ct2 = roundtrip(kb, ctx, kb.WIRE_CIPHERTEXT, ct)
    dec = ctx.decrypt(sk, ct2)
    ctx.decode(dec)
    assert np.array_equal(kb.to_host(dec), m)
    roundtrip(kb, ctx, kb.WIRE_PLAINTEXT, kb.to_device(m))
    roundtrip(kb, ctx, kb.WIRE_SECRET_KEY, sk[None])
    roundtrip(kb, ctx, kb.WIRE_PUBLIC_KEY, pk[None])
    rlk = ctx.keygen_rlk(bytes([4]) * 32, sk)
    roundtrip(kb, ctx, kb.WIRE_RELIN_KEY, rlk[None])
    g = ctx.galois_elt(1)
    gk = ctx.keygen_galois(bytes([5]) * 32, sk, g)
    roundtrip(kb, ctx, kb.WIRE_GALOIS_KEY, gk[None], aux=g)
    mg = np.zeros(ctx.n, dtype=np.uint64)
    mg[1] = 1
    rg = ctx.rgsw_encrypt(bytes([6]) * 32, sk, kb.to_device(mg[None])[0], 20)
    roundtrip(kb, ctx, kb.WIRE_RGSW, rg[None], aux=20)


def _reject(kb, fn, msg):
    with pytest.raises(kb.KbfvError) as e:
        fn()
    assert msg in str(e.value), str(e.value)


def test_loader_rejections(setup):
    kb, ctx, sk, pk, m, ct = setup
    blob = ctx.wire_save(kb.WIRE_CIPHERTEXT, ct)
    out = ctx.empty(*ct.shape)
    _reject(kb, lambda: ctx.wire_load(blob, kb.WIRE_PUBLIC_KEY, out), "kind")
    _reject(kb, lambda: ctx.wire_load(blob, kb.WIRE_CIPHERTEXT, ctx.empty(2, 2, ctx.qn, ctx.n)), "objects")
    _reject(kb, lambda: ctx.wire_load(blob[:-8], kb.WIRE_CIPHERTEXT, out), "truncated")
    bad = bytearray(blob)
    bad[64 + 8 * 100] ^= 0x10
    _reject(kb, lambda: ctx.wire_load(bytes(bad), kb.WIRE_CIPHERTEXT, out), "checksum")
    # a residue equal to its limb's modulus, checksum made consistent
    words = np.frombuffer(blob[64:], dtype=np.uint64).copy()
    words[ctx.n * 1 + 7] = np.uint64(ctx.mod[ctx.oQ + 1])  # batch 0, poly 0, limb 1
    head = bytearray(blob[:64])
    head[56:64] = struct.pack("<Q", kb.wire_checksum(words))
    _reject(kb, lambda: ctx.wire_load(bytes(head) + words.tobytes(), kb.WIRE_CIPHERTEXT, out), "range")
    # other parameters (same shape, different chain)
    d = P.chain60(12, 3, 1, 3, 1, plain=257)
    other = kb.Context(d["log_n"], d["q"], d["p"], d["t"], d["gamma"], d["plain"])
    _reject(kb, lambda: other.wire_load(blob, kb.WIRE_CIPHERTEXT, other.empty(*ct.shape)), "parameters")
    # the context still loads the untouched blob after every rejection
    ctx.wire_load(blob, kb.WIRE_CIPHERTEXT, out)
    assert np.array_equal(kb.to_host(out), kb.to_host(ct))


def test_saver_rejects_out_of_range_payload(setup):
    """What the loader would refuse is refused at save time: a plaintext in the
    NTT domain (residues mod q, not coefficients mod t, the encode_ntt form)
    and a ciphertext with a residue equal to its limb's modulus."""
    kb, ctx, sk, pk, m, ct = setup
    pt = kb.to_device(m[:1].copy())
    ctx.encode(pt)
    ctx.wire_save(kb.WIRE_PLAINTEXT, pt)  # the encode() form saves
    lifted = kb.to_device(np.broadcast_to(kb.to_host(pt)[0], (ctx.qn, ctx.n)).copy())
    ctx.ntt(lifted, lifted, +1, ctx.oQ, ctx.qn)  # limb 0 of the encode_ntt form
    ntt_form = lifted[0:1].contiguous()
    assert (kb.to_host(ntt_form) >= ctx.plain).any()
    _reject(kb, lambda: ctx.wire_save(kb.WIRE_PLAINTEXT, ntt_form), "encode()")
    bad = kb.to_host(ct).copy()
    bad[0, 0, 1, 7] = ctx.mod[ctx.oQ + 1]
    _reject(kb, lambda: ctx.wire_save(kb.WIRE_CIPHERTEXT, kb.to_device(bad)), "range")
