"""NTT plan space (host only, no device calls): the candidate list holds the
built-in plan's shape and exactly the (split, tile, grid order) combinations
whose passes the kernels cover, and the validator rejects malformed words."""
import os

import pytest


def expect_pass_ok(log_n, log_s, log_t, column):
    # 2^log_s-point sub-transforms, tiles of 2^log_t elements (16 values per
    # thread): two register rounds (radix-32 only at 2^9), at least one tile
    # per limb and, for the column pass, rows of at least 8 columns (64 bytes)
    if not (5 <= log_s <= 9 and 11 <= log_t <= 13):
        return False
    if log_t < log_s or log_t > log_n:
        return False
    return not (column and log_t - log_s < 3)


@pytest.mark.parametrize("log_n", [12, 13, 14, 15, 16])
def test_candidates_are_exactly_the_valid_shapes(log_n):
    import kunpbfv_amd as kb
    got = kb.ntt_plan_candidates(log_n)
    want = set()
    if log_n <= 14:  # single pass: the whole limb in one workgroup's LDS
        want |= {log_n | cx << 12 for cx in (0, 1)}
    c0 = (log_n + 1) // 2
    for cs in (c0 - 1, c0, c0 + 1):
        for ct in (11, 12, 13):
            for rt in (11, 12, 13):
                if expect_pass_ok(log_n, cs, ct, True) and expect_pass_ok(log_n, log_n - cs, rt, False):
                    for cx in (0, 1):
                        for rx in (0, 1):
                            want.add(cs | ct << 4 | rt << 8 | cx << 12 | rx << 13)
    assert set(got) == want and len(got) == len(want)
    for p in got:
        assert kb.lib.kbfv_ntt_plan_valid(log_n, p) == 0
        f = kb.ntt_plan_fields(p, log_n)
        assert f[0] + (log_n - f[0]) == log_n
        if f[0] == log_n:
            assert f[1:3] == (0, 0) and f[4] == 0
    # the built-in plan (0) is the balanced split, 4096-element tiles, XCD order
    assert kb.lib.kbfv_ntt_plan_valid(log_n, 0) == 0
    builtin = c0 | 12 << 4 | 12 << 8 | 1 << 12 | 1 << 13
    assert builtin in got and kb.ntt_plan_fields(0, log_n) == kb.ntt_plan_fields(builtin, log_n)


def test_large_rings_offer_radix32_splits():
    import kunpbfv_amd as kb
    splits = {p & 15 for p in kb.ntt_plan_candidates(16)}
    assert splits == {7, 8, 9}
    # a limb of 2^15 (256 KiB) does not fit one workgroup's LDS: no single pass
    assert 15 not in {p & 15 for p in kb.ntt_plan_candidates(15)}
    assert 14 in {p & 15 for p in kb.ntt_plan_candidates(14)}
    assert kb.lib.kbfv_ntt_plan_valid(15, 15) != 0
    assert kb.lib.kbfv_ntt_plan_valid(14, 14 | 12 << 4) != 0   # a single pass has no tiles
    # a 2^9-point column pass needs 8 columns of 512: tiles of 2^12 or more
    assert not any((p & 15) == 9 and ((p >> 4) & 15) == 11 for p in kb.ntt_plan_candidates(16))


def test_validator_rejects_malformed_words():
    import kunpbfv_amd as kb
    lib = kb.lib
    assert lib.kbfv_ntt_plan_valid(16, 1 << 17 | 8 | 12 << 4 | 12 << 8) != 0   # unknown bit
    assert lib.kbfv_ntt_plan_valid(16, 1 << 14 | 8 | 12 << 4 | 12 << 8) != 0   # chunk code 1 (8 MB) unused
    assert lib.kbfv_ntt_plan_valid(16, 7 << 14 | 8 | 12 << 4 | 12 << 8) != 0   # chunk code 7 unused
    assert lib.kbfv_ntt_plan_valid(16, 10 | 13 << 4 | 13 << 8) != 0           # split too far off
    assert lib.kbfv_ntt_plan_valid(12, 6 | 13 << 4 | 12 << 8) != 0            # tile larger than the ring
    assert lib.kbfv_ntt_plan_valid(11, 0) != 0                                 # ring degree
    assert [lib.kbfv_ntt_plan_bucket(n) for n in (1, 63, 64, 511, 512, 4095, 4096, 1 << 20)] == \
        [0, 0, 1, 1, 2, 2, 3, 3]


def test_plan_cache_file_round_trip(tmp_path):
    """The JSON plan cache (host only): entries merge into one file, a missing
    or corrupt file reads as empty, and a write replaces the file whole."""
    import json
    import kunpbfv_amd as kb
    path = str(tmp_path / "sub" / "plans.json")
    assert kb._cache_read(path) == {}
    kb._cache_write(path, "a", [[1, 0, 0x3cd8, 0.04, 57]])
    kb._cache_write(path, "b", [[-1, 3, 0x1cd8, 11.2, 21888]])
    data = kb._cache_read(path)
    assert data == {"a": [[1, 0, 0x3cd8, 0.04, 57]], "b": [[-1, 3, 0x1cd8, 11.2, 21888]]}
    assert json.load(open(path)) == data
    kb._cache_write(path, "a", [[1, 0, 0, 0.05, 57]])
    assert kb._cache_read(path)["a"] == [[1, 0, 0, 0.05, 57]]
    open(path, "w").write("{not json")
    assert kb._cache_read(path) == {}
    kb._cache_write(path, "c", [])
    assert kb._cache_read(path) == {"c": []}
    assert not [f for f in (tmp_path / "sub").iterdir() if ".tmp" in f.name]


def test_chunked_plans_are_valid_two_pass_plans():
    """Chunk codes 2..6 (both passes per chunk of 2^(k+23) bytes) are valid on
    every two-pass candidate and on none of the single-pass plans."""
    import kunpbfv_amd as kb
    for log_n in (12, 14, 16):
        for p in kb.ntt_plan_candidates(log_n):
            single = (p & 15) == log_n
            for k in range(2, 7):
                ok = kb.lib.kbfv_ntt_plan_valid(log_n, p | k << 14) == 0
                assert ok != single, (log_n, hex(p), k)
                assert kb.ntt_plan_chunk_bytes(p | k << 14) == 1 << (k + 23)
            assert kb.ntt_plan_chunk_bytes(p) == 0


def test_plan_cache_concurrent_writers_keep_every_entry(tmp_path):
    """Processes adding entries at once lose none of them (the cache is updated
    under a file lock), and the cache key names the library build."""
    import subprocess
    import sys
    import kunpbfv_amd as kb
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    path = str(tmp_path / "plans.json")
    code = ("import sys, kunpbfv_amd as kb\n"
            "for i in range(40): kb._cache_write(sys.argv[1], f'{sys.argv[2]}-{i}', [i])\n")
    procs = [subprocess.Popen([sys.executable, "-c", code, path, f"w{j}"], cwd=root) for j in range(4)]
    assert all(p.wait() == 0 for p in procs)
    data = kb._cache_read(path)
    assert set(data) == {f"w{j}-{i}" for j in range(4) for i in range(40)}
    bid = kb.library_build_id()
    assert len(bid) == 16 and bid == kb.library_build_id()
