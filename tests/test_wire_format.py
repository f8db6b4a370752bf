"""The wire format's host side (CPU only): the payload checksum against its
definition restated in numpy, and header parsing of blobs built byte by byte
from the specification (magic, version, kind, count, aux, payload size)."""
import struct

import numpy as np
import pytest

M64 = (1 << 64) - 1


def checksum_restated(words):
    """s0 ^ rotl(s1, 32), s0 = sum w_i, s1 = sum (i + 1) w_i (mod 2^64)."""
    w = np.asarray(words, dtype=np.uint64)
    idx = np.arange(1, w.size + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        s0 = int(w.sum(dtype=np.uint64))
        s1 = int((w * idx).sum(dtype=np.uint64))
    return s0 ^ (((s1 << 32) | (s1 >> 32)) & M64)


def blob(kind=1, count=2, aux=0, log_n=12, dims=(2, 3, 1, 1), payload_words=None, magic=b"KBFV", version=1):
    words = np.zeros(0, np.uint64) if payload_words is None else np.asarray(payload_words, np.uint64)
    head = magic + struct.pack("<HHII4IQQQQ", version, kind, log_n, count, *dims, aux, 0x1234, words.nbytes,
                               checksum_restated(words))
    assert len(head) == 64
    return head + words.tobytes()


def test_checksum_matches_definition():
    import kunpbfv_amd as kb
    rng = np.random.default_rng(7)
    cases = [np.zeros(0, np.uint64), np.array([0], np.uint64), np.array([M64], np.uint64),
             np.full(1000, M64, np.uint64), rng.integers(0, 1 << 63, size=4097, dtype=np.uint64) * 2 + 1]
    for w in cases:
        assert kb.wire_checksum(w) == checksum_restated(w)
    # order-sensitive: swapping two words changes it, a flipped bit changes it
    w = rng.integers(0, 1 << 60, size=64, dtype=np.uint64)
    sw = w.copy()
    sw[[3, 9]] = sw[[9, 3]]
    assert kb.wire_checksum(sw) != kb.wire_checksum(w)
    fl = w.copy()
    fl[17] ^= np.uint64(1 << 41)
    assert kb.wire_checksum(fl) != kb.wire_checksum(w)


def test_peek_reads_the_header_fields():
    import kunpbfv_amd as kb
    b = blob(kind=6, count=3, aux=12345, payload_words=np.arange(10))
    assert kb.wire_peek(b) == (6, 3, 12345)
    # extra trailing bytes are allowed; the payload size field decides
    assert kb.wire_peek(b + b"\0" * 8) == (6, 3, 12345)


@pytest.mark.parametrize("bad, msg", [
    (lambda: blob(magic=b"KBFW"), b"magic"),
    (lambda: blob(version=2), b"version"),
    (lambda: blob(payload_words=np.arange(8))[:-8], b"truncated"),
    (lambda: blob()[:40], b"header"),
])
def test_peek_rejects_malformed_blobs(bad, msg):
    import kunpbfv_amd as kb
    b = bad()
    with pytest.raises(kb.KbfvError) as e:
        kb.wire_peek(b)
    assert msg.decode() in str(e.value)
