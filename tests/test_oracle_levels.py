"""Oracle pins for ciphertexts below the full level (CPU only).

* or_mod_switch (the reference's rescaleToAndEqual, include/BFVScheme.cuh:
  469-487) against its closed form: with X the CRT value in [0, Q_from) of a
  coefficient, the result is floor(X / (q_(to+1) ... q_from)) mod q_i, the
  composition of the per-limb steps floor(x / q_l); checked with Python big
  integers on every coefficient of a small ring.  A planted mistake (the
  centred residue of x_l, i.e. a rounding switch) fails it
  (test_oracle_mutations.py).
* or_ctx_level: the level context at the top level equals the full context
  table for table and HomMult for HomMult; below it, its R basis is the
  level + 2 moduli after q_level in the full [Q | R] table and every HPS table
  equals its closed form (Python big integers over Q_l, R_l).
* end to end at every level: an encryption switched down decrypts to its
  message, and HomMult + relinearisation with the full-level key at the level
  decrypts to the slot-wise product, for one and two special primes (K = 2
  makes the last digit group shorter than at the full level).
"""
import numpy as np
import pytest

import oracle_lib as orc
import params as P


def crt(residues, mods):
    M = 1
    for m in mods:
        M *= m
    x = 0
    for r, m in zip(residues, mods):
        Mi = M // m
        x += int(r) * Mi * pow(Mi, -1, m)
    return x % M


def _setup(d, seed=1):
    oc = orc.Ctx(d["log_n"], d["q"], d["p"], d["t"], d["gamma"], d["plain"])
    sk = oc.keygen_sk(bytes([seed]) * 32)
    pk = oc.keygen_pk(bytes([seed + 1]) * 32, sk)
    return oc, sk, pk


SETS = {"k1_n4096": P.chain60(12, 4, 1, 4, 2), "k2_ref14": P.REF14}


def check_mod_switch_floor(oc, frm, to):
    q = [int(v) for v in oc.mod[oc.oQ:oc.oQ + oc.qn]]
    rng = np.random.default_rng(frm * 10 + to)
    x = np.empty((2, frm + 1, oc.n), dtype=np.uint64)
    for i in range(frm + 1):
        x[:, i] = rng.integers(0, q[i], size=(2, oc.n), dtype=np.uint64)
    x[:, :, 0] = 0                               # X = 0
    for i in range(frm + 1):
        x[:, i, 1] = q[i] - 1                    # X = -1 mod Q_from (largest value)
    xn = oc.ntt_fwd(x, oc.oQ, frm + 1, polys=2)
    got = oc.ntt_inv(oc.mod_switch(xn, frm, to), oc.oQ, to + 1, polys=2)
    div = 1
    for l in range(to + 1, frm + 1):
        div *= q[l]
    for h in range(2):
        for k in range(oc.n):
            X = crt(x[h, :, k], q[:frm + 1])
            want = X // div
            for i in range(to + 1):
                assert int(got[h, i, k]) == want % q[i], (h, k, i)


def ctx_k1_n4096(lib=None):
    d = P.chain60(12, 4, 1, 4, 2)
    return orc.Ctx(d["log_n"], d["q"], d["p"], d["t"], d["gamma"], d["plain"], _lib=lib)


@pytest.mark.parametrize("frm,to", [(3, 0), (3, 2), (2, 1)])
def test_mod_switch_is_floor_division(frm, to):
    check_mod_switch_floor(ctx_k1_n4096(), frm, to)


def test_top_level_context_equals_full():
    d = P.chain60(12, 3, 1, 4, 2)
    oc, sk, pk = _setup(d)
    top = oc.level(oc.qn - 1)
    for name in ("q_hatinv", "q_hat_mod_r", "alphaQ_mod_r", "r_hatinv", "r_hat_mod_q", "alphaR_mod_q",
                 "sr_modr", "dec_modt", "dec_B_modt", "qj_hatinv", "ri_hat_mod_t", "t_hat_mod_pq"):
        assert np.array_equal(top.export(name), oc.export(name)), name
    assert np.array_equal(top.mod, oc.mod)
    rlk = oc.keygen_rlk(bytes([9]) * 32, sk)
    rng = np.random.default_rng(3)
    m = [oc.encode(rng.integers(0, d["plain"], size=oc.n, dtype=np.uint64)) for _ in range(2)]
    c1, c2 = (oc.encrypt(bytes([4]) * 32, pk, mm, entry=e) for e, mm in enumerate(m))
    assert np.array_equal(top.mult_relin(rlk, c1, c2), oc.mult_relin(rlk, c1, c2))


def test_level_hps_tables_closed_forms():
    d = P.chain60(12, 5, 1, 4, 2)
    oc = orc.Ctx(d["log_n"], d["q"], d["p"], d["t"], d["gamma"], d["plain"])
    for lvl in (0, 2, 3):
        lc = oc.level(lvl)
        assert (lc.qn, lc.rn) == (lvl + 1, lvl + 2)
        Q = [int(v) for v in oc.mod[oc.oQ:oc.oQ + lvl + 1]]
        R = [int(v) for v in oc.mod[oc.oQ + lvl + 1:oc.oQ + 2 * lvl + 3]]
        assert [int(v) for v in lc.mod[lc.oQ:lc.oQ + lc.qn]] == Q
        assert [int(v) for v in lc.mod[lc.oR:lc.oR + lc.rn]] == R
        Qp = 1
        for v in Q:
            Qp *= v
        assert [int(v) for v in lc.export("q_hatinv")] == [pow(Qp // v, -1, v) for v in Q]
        qhr = lc.export("q_hat_mod_r").reshape(len(R), len(Q))
        for j, r in enumerate(R):
            assert [int(v) for v in qhr[j]] == [(Qp // v) % r for v in Q]
        # scale-and-round: floor(t R (S/s_i)^-1 / s_i) mod r_j for the Q part
        S = Q + R
        Sp = Qp
        for v in R:
            Sp *= v
        Rp = Sp // Qp
        srm = lc.export("sr_modr").reshape(len(R), len(Q) + 1)
        for i, s in enumerate(Q):
            f = (d["plain"] * Rp * pow(Sp // s, -1, s)) // s
            assert [int(srm[j, i]) for j in range(len(R))] == [f % r for r in R], (lvl, i)


@pytest.mark.parametrize("name", list(SETS))
def test_switched_down_ciphertexts_decrypt_and_multiply(name):
    d = SETS[name]
    oc, sk, pk = _setup(d, seed=5)
    rlk = oc.keygen_rlk(bytes([8]) * 32, sk)
    rng = np.random.default_rng(11)
    s1, s2 = (rng.integers(0, d["plain"], size=oc.n, dtype=np.uint64) for _ in range(2))
    c1 = oc.encrypt(bytes([6]) * 32, pk, oc.encode(s1))
    c2 = oc.encrypt(bytes([6]) * 32, pk, oc.encode(s2), entry=1)
    want = (s1.astype(object) * s2.astype(object) % d["plain"]).astype(np.uint64)
    for lvl in range(oc.qn - 1, 0, -1):
        lc = oc.level(lvl)
        sk_l = np.ascontiguousarray(sk[:oc.pn + lvl + 1])
        a = oc.mod_switch(c1, oc.qn - 1, lvl)
        b = oc.mod_switch(c2, oc.qn - 1, lvl)
        assert np.array_equal(lc.decode(lc.decrypt(sk_l, a)), s1), (name, lvl)
        prod = lc.mult_relin(rlk, a, b)
        assert np.array_equal(lc.decode(lc.decrypt(sk_l, prod)), want), (name, lvl)
