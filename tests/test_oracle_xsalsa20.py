"""Pins the oracle's XSalsa20 and samplers (CPU only): published NaCl test
vectors, an independent pure-Python XSalsa20 written from Bernstein's
specification, and distribution tests for the samplers."""
import random
import struct

import numpy as np
import pytest

import oracle_lib as orc

# NaCl test vectors (tests/core1.c and tests/stream3.c of the NaCl distribution)
SHARED = bytes([0x4a, 0x5d, 0x9d, 0x5b, 0xa4, 0xce, 0x2d, 0xe1, 0x72, 0x8e, 0x3b, 0xf4, 0x80, 0x35, 0x0f, 0x25,
                0xe0, 0x7e, 0x21, 0xc9, 0x47, 0xd1, 0x9e, 0x33, 0x76, 0xf0, 0x9b, 0x3c, 0x1e, 0x16, 0x17, 0x42])
FIRSTKEY = bytes([0x1b, 0x27, 0x55, 0x64, 0x73, 0xe9, 0x85, 0xd4, 0x62, 0xcd, 0x51, 0x19, 0x7a, 0x9a, 0x46, 0xc7,
                  0x60, 0x09, 0x54, 0x9e, 0xac, 0x64, 0x74, 0xf2, 0x06, 0xc4, 0xee, 0x08, 0x44, 0xf6, 0x83, 0x89])
NONCE = bytes([0x69, 0x69, 0x6e, 0xe9, 0x55, 0xb6, 0x2b, 0x73, 0xcd, 0x62, 0xbd, 0xa8, 0x75, 0xfc, 0x73, 0xd6,
               0x82, 0x19, 0xe0, 0x03, 0x6b, 0x7a, 0x0b, 0x37])
STREAM32 = bytes([0xee, 0xa6, 0xa7, 0x25, 0x1c, 0x1e, 0x72, 0x91, 0x6d, 0x11, 0xc2, 0xcb, 0x21, 0x4d, 0x3c, 0x25,
                  0x25, 0x39, 0x12, 0x1d, 0x8e, 0x23, 0x4e, 0x65, 0x2d, 0x65, 0x1f, 0xa4, 0xc8, 0xcf, 0xf8, 0x80])


def _rotl(x, c):
    return ((x << c) | (x >> (32 - c))) & 0xffffffff


def _quarter(y0, y1, y2, y3):
    y1 ^= _rotl((y0 + y3) & 0xffffffff, 7)
    y2 ^= _rotl((y1 + y0) & 0xffffffff, 9)
    y3 ^= _rotl((y2 + y1) & 0xffffffff, 13)
    y0 ^= _rotl((y3 + y2) & 0xffffffff, 18)
    return y0, y1, y2, y3


def _doubleround(x):
    # columnround then rowround, as the specification's quarterround lists them
    x = list(x)
    for a, b, c, d in ((0, 4, 8, 12), (5, 9, 13, 1), (10, 14, 2, 6), (15, 3, 7, 11)):
        x[a], x[b], x[c], x[d] = _quarter(x[a], x[b], x[c], x[d])
    for a, b, c, d in ((0, 1, 2, 3), (5, 6, 7, 4), (10, 11, 8, 9), (15, 12, 13, 14)):
        x[a], x[b], x[c], x[d] = _quarter(x[a], x[b], x[c], x[d])
    return x


SIG = struct.unpack("<4I", b"expand 32-byte k")


def py_hsalsa20(key, n16):
    k = struct.unpack("<8I", key)
    n = struct.unpack("<4I", n16)
    x = [SIG[0], *k[:4], SIG[1], *n, SIG[2], *k[4:], SIG[3]]
    for _ in range(10):
        x = _doubleround(x)
    return struct.pack("<8I", x[0], x[5], x[10], x[15], x[6], x[7], x[8], x[9])


def py_xsalsa20(key, nonce, length):
    sub = py_hsalsa20(key, nonce[:16])
    k = struct.unpack("<8I", sub)
    n = struct.unpack("<2I", nonce[16:])
    out = b""
    ctr = 0
    while len(out) < length:
        x0 = [SIG[0], *k[:4], SIG[1], n[0], n[1], ctr & 0xffffffff, ctr >> 32, SIG[2], *k[4:], SIG[3]]
        x = x0
        for _ in range(10):
            x = _doubleround(x)
        out += struct.pack("<16I", *[(a + b) & 0xffffffff for a, b in zip(x, x0)])
        ctr += 1
    return out[:length]


def test_nacl_vectors():
    # HSalsa20(shared secret, zero nonce) = firstkey (core1.c)
    assert py_hsalsa20(SHARED, bytes(16)) == FIRSTKEY
    assert orc.xsalsa20(SHARED, bytes(24), 0) == b""  # length-0 stream is empty
    assert orc.xsalsa20(FIRSTKEY, NONCE, 32) == STREAM32
    assert py_xsalsa20(FIRSTKEY, NONCE, 32) == STREAM32


def test_oracle_matches_spec_implementation_on_random_keys():
    rng = random.Random(5)
    for length in (1, 63, 64, 65, 300):
        key = bytes(rng.randrange(256) for _ in range(32))
        nonce = bytes(rng.randrange(256) for _ in range(24))
        assert orc.xsalsa20(key, nonce, length) == py_xsalsa20(key, nonce, length)


def test_block_offset():
    s = orc.xsalsa20(FIRSTKEY, NONCE, 64 * 6)
    assert orc.xsalsa20(FIRSTKEY, NONCE, 128, block0=4) == s[256:384]


@pytest.fixture(scope="module")
def oc():
    q, p, t = orc.chain_60bit(2, 1, 3)
    return orc.Ctx(13, q, p, t, 1, 65537)


def _stream_words(seed, tag, entry, count, width):
    nonce = struct.pack("<Q", tag) + bytes(8) + struct.pack("<Q", entry)
    raw = py_xsalsa20(seed, nonce, count * width)
    return raw


def test_sampler_layout_follows_the_stream(oc):
    seed = bytes(range(32))
    tag = (7 << 32) | 3
    n = oc.n
    raw = _stream_words(seed, tag, 2, n, 8)
    words = np.frombuffer(raw, dtype="<u8")
    tern = oc.sample("ternary", seed, tag, oc.oQ, 1, entry=2)[0]
    q0 = int(oc.mod[oc.oQ])
    want = [(int(w) % 3) - 1 for w in words]
    assert [int(v) if v < q0 // 2 else int(v) - q0 for v in tern] == want
    cbd = oc.sample("cbd", seed, tag, oc.oQ, 1, entry=2)[0]
    M = (1 << 21) - 1
    want = [bin(int(w) & M).count("1") - bin((int(w) >> 21) & M).count("1") for w in words]
    assert [int(v) if v < q0 // 2 else int(v) - q0 for v in cbd] == want
    raw16 = np.frombuffer(_stream_words(seed, tag, 0, n, 16), dtype="<u8").reshape(-1, 2)
    uni = oc.sample("uniform", seed, tag, oc.oQ, 1)[0]
    assert [int(v) for v in uni] == [(int(hi) << 64 | int(lo)) % q0 for lo, hi in raw16]


def test_sampler_distributions(oc):
    seed = bytes([9]) * 32
    q0 = int(oc.mod[oc.oQ])
    sig = lambda v: v.astype(np.int64) - np.where(v > q0 // 2, q0, 0)  # noqa: E731
    t = sig(oc.sample("ternary", seed, 1 << 32, oc.oQ, oc.qn)[0])
    counts = np.bincount(t + 1, minlength=3)
    assert set(np.unique(t)) <= {-1, 0, 1}
    assert np.all(np.abs(counts - oc.n / 3) < 5 * np.sqrt(oc.n))
    z = sig(oc.sample("zo", seed, 2 << 32, oc.oQ, oc.qn)[0])
    assert abs((z == 0).mean() - 0.5) < 0.03 and abs((z == 1).mean() - 0.25) < 0.03
    c = sig(oc.sample("cbd", seed, 3 << 32, oc.oQ, oc.qn)[0]).astype(float)
    assert abs(c.mean()) < 0.2 and abs(c.var() - 10.5) < 1.0 and np.abs(c).max() <= 21
    u = oc.sample("uniform", seed, 4 << 32, oc.oQ, oc.qn)
    for i in range(oc.qn):
        q = int(oc.mod[oc.oQ + i])
        hist = np.histogram(u[i].astype(float) / q, bins=16, range=(0, 1))[0]
        chi2 = ((hist - oc.n / 16) ** 2 / (oc.n / 16)).sum()
        assert chi2 < 50  # 15 degrees of freedom
    # every limb carries the same small coefficient
    c_all = oc.sample("cbd", seed, 3 << 32, oc.oQ, oc.qn)
    for i in range(1, oc.qn):
        qi = int(oc.mod[oc.oQ + i])
        assert np.array_equal(c_all[i].astype(np.int64) - np.where(c_all[i] > qi // 2, qi, 0), c)
