"""Pins the oracle's HPS pieces with exact CRT arithmetic in Python big
integers (CPU only).  The reference's double-precision estimates round a sum
whose exact value is known here; the two agree except when that sum sits
within floating-point error of a rounding tie, which these inputs exclude."""
import numpy as np
import pytest

import oracle_lib as orc


def ctx_small():
    q, p, t = orc.chain_60bit(4, 1, 3)
    return orc.Ctx(12, q, p, t, 1, 65537)


@pytest.fixture(scope="module")
def oc():
    return ctx_small()


def crt(residues, mods):
    M = 1
    for m in mods:
        M *= m
    x = 0
    for r, m in zip(residues, mods):
        Mi = M // m
        x += int(r) * Mi * pow(Mi, -1, m)
    return x % M, M


def near_tie(v, scale, eps=2.0 ** -30):
    """v/scale within eps of a half-integer"""
    from fractions import Fraction
    f = Fraction(v, scale) % 1
    return abs(f - Fraction(1, 2)) < eps


def test_bconv_q_to_r_is_centered_lift(oc):
    rng = np.random.default_rng(1)
    Q = [int(x) for x in oc.mod[oc.oQ:oc.oQ + oc.qn]]
    R = [int(x) for x in oc.mod[oc.oR:oc.oR + oc.rn]]
    x = np.stack([rng.integers(0, q, size=oc.n, dtype=np.uint64) for q in Q])
    y = oc.bconv_q_to_r(x)
    checked = 0
    for k in range(0, oc.n, 37):
        X, Qp = crt(x[:, k], Q)
        if near_tie(X, Qp):
            continue
        Xc = X - Qp if 2 * X >= Qp else X
        assert [int(v) for v in y[:, k]] == [Xc % r for r in R]
        checked += 1
    assert checked > 100


def test_bconv_r_to_q_is_centered_lift(oc):
    rng = np.random.default_rng(2)
    Q = [int(x) for x in oc.mod[oc.oQ:oc.oQ + oc.qn]]
    R = [int(x) for x in oc.mod[oc.oR:oc.oR + oc.rn]]
    x = np.stack([rng.integers(0, r, size=oc.n, dtype=np.uint64) for r in R])
    y = oc.bconv_r_to_q(x)
    for k in range(0, oc.n, 41):
        X, Rp = crt(x[:, k], R)
        if near_tie(X, Rp):
            continue
        Xc = X - Rp if 2 * X >= Rp else X
        assert [int(v) for v in y[:, k]] == [Xc % q for q in Q]


def check_scale_round_exact_small(oc, seed=13):
    """Q residues below 2^20 and arbitrary R residues: the double estimate of
    sum_i x_i * frac_i is then exact to about 2^-27 (each term errs by at most
    x_i * 2^-54, each fma by half an ulp of a sum below 2^25), so away from a
    rounding tie the result must be exactly floor(t X / Q + 1/2) mod r_j, X the
    CRT value over Q u R.  A dropped +1/2, or nu rounded instead of truncated
    (include/poly_arithmetic.cuh:1276-1290), moves about half of the outputs."""
    from fractions import Fraction
    rng = np.random.default_rng(seed)
    S = [int(x) for x in oc.mod[oc.oQ:oc.oQ + oc.qn + oc.rn]]
    R = S[oc.qn:]
    Qp = 1
    for q in S[:oc.qn]:
        Qp *= q
    x = np.stack([rng.integers(0, 1 << 20, size=oc.n, dtype=np.uint64) for _ in range(oc.qn)]
                 + [rng.integers(0, r, size=oc.n, dtype=np.uint64) for r in R])
    x[:oc.qn, :8] = 0  # X = the R part alone: t X / Q integral or far from a tie
    y = oc.scale_round(x)
    t = oc.plain
    checked = 0
    for k in range(0, oc.n, 3):
        X, _ = crt(x[:, k], S)
        f = Fraction(t * X, Qp) % 1
        if abs(f - Fraction(1, 2)) < Fraction(1, 1 << 20):
            continue
        want = (2 * t * X + Qp) // (2 * Qp)
        got = [int(v) for v in y[:, k]]
        assert got == [want % r for r in R], (k, got, [want % r for r in R])
        checked += 1
    assert checked > 1000
    return checked


def test_scale_round_exact_on_small_residues(oc):
    check_scale_round_exact_small(oc)


def test_scale_round_is_round_t_x_over_q(oc):
    """Full-size residues: floor(t X / Q + 1/2) in every R prime, up to one
    small offset d common to all R primes.  The reference estimates the
    rounding term from 60-bit residues in double precision
    (include/poly_arithmetic.cuh:1276-1282); a product x_i * frac_i near 2^60
    carries rounding errors of a few hundred units, which HPS absorbs as noise
    (the exact case is test_scale_round_exact_on_small_residues)."""
    rng = np.random.default_rng(3)
    S = [int(x) for x in oc.mod[oc.oQ:oc.oQ + oc.qn + oc.rn]]
    R = S[oc.qn:]
    Qp = 1
    for q in S[:oc.qn]:
        Qp *= q
    x = np.stack([rng.integers(0, s, size=oc.n, dtype=np.uint64) for s in S])
    y = oc.scale_round(x)
    t = oc.plain
    worst = 0
    for k in range(0, oc.n, 53):
        X, _ = crt(x[:, k], S)
        want = (2 * t * X + Qp) // (2 * Qp)  # floor(t X / Q + 1/2)
        offs = {(int(y[j, k]) - want) % r for j, r in enumerate(R)}
        d = [(int(y[j, k]) - want) % r for j, r in enumerate(R)]
        d = [v - r if v > r // 2 else v for v, r in zip(d, R)]
        assert len(set(d)) == 1, offs  # the same integer offset in every R prime
        worst = max(worst, abs(d[0]))
    assert worst < 1 << 12


def test_hps_decrypt_recovers_messages(oc):
    """Inputs shaped like decryptable values X = round(Q m / t) + e, |e| < Q/(4t):
    the message comes back exactly."""
    rng = np.random.default_rng(4)
    Q = [int(x) for x in oc.mod[oc.oQ:oc.oQ + oc.qn]]
    Qp = 1
    for q in Q:
        Qp *= q
    t = oc.plain
    m = rng.integers(0, t, size=oc.n)
    e = [int(v) for v in rng.integers(-(1 << 62), 1 << 62, size=oc.n)]
    X = [((Qp * int(mi) + t // 2) // t + ei * (Qp >> 130)) % Qp for mi, ei in zip(m, e)]
    x = np.array([[v % q for v in X] for q in Q], dtype=np.uint64)
    assert np.array_equal(oc.hps_decrypt(x), m.astype(np.uint64))


def test_hps_decrypt_exact_on_small_residues(oc):
    """Residues below 2^20: the double sums are exact to far below 2^-20, so
    the result is round(t X / Q) mod t exactly away from a tie."""
    from fractions import Fraction
    rng = np.random.default_rng(14)
    Q = [int(x) for x in oc.mod[oc.oQ:oc.oQ + oc.qn]]
    x = np.stack([rng.integers(0, 1 << 20, size=oc.n, dtype=np.uint64) for _ in Q])
    m = oc.hps_decrypt(x)
    t = oc.plain
    checked = 0
    for k in range(0, oc.n, 3):
        X, Qp = crt(x[:, k], Q)
        f = Fraction(t * X, Qp) % 1
        if abs(f - Fraction(1, 2)) < Fraction(1, 1 << 20):
            continue
        assert int(m[k]) % t == ((2 * t * X + Qp) // (2 * Qp)) % t, k
        checked += 1
    assert checked > 1000


def test_hps_decrypt_uniform_inputs_round_within_estimate(oc):
    """On arbitrary inputs the result is round(t X / Q) mod t whenever the
    fractional part is further than 0.3 from 1/2 (the lazy variant adds a
    ~2^50 integer sum into the double before rounding,
    include/poly_arithmetic.cuh:1068-1072), and otherwise one of its two
    neighbours."""
    rng = np.random.default_rng(4)
    Q = [int(x) for x in oc.mod[oc.oQ:oc.oQ + oc.qn]]
    x = np.stack([rng.integers(0, q, size=oc.n, dtype=np.uint64) for q in Q])
    m = oc.hps_decrypt(x)
    t = oc.plain
    from fractions import Fraction
    for k in range(0, oc.n, 29):
        X, Qp = crt(x[:, k], Q)
        fl = (t * X) // Qp
        frac = Fraction(t * X, Qp) - fl
        got = int(m[k]) % t
        if abs(frac - Fraction(1, 2)) > Fraction(3, 10):
            assert got == ((2 * t * X + Qp) // (2 * Qp)) % t
        else:
            assert got in (fl % t, (fl + 1) % t)


def test_tensor_is_mod_q_products(oc):
    rng = np.random.default_rng(5)
    Q = [int(x) for x in oc.mod[oc.oQ:oc.oQ + oc.qn]]
    c1 = np.stack([np.stack([rng.integers(0, q, size=oc.n, dtype=np.uint64) for q in Q]) for _ in range(2)])
    c2 = np.stack([np.stack([rng.integers(0, q, size=oc.n, dtype=np.uint64) for q in Q]) for _ in range(2)])
    out = oc.tensor(c1, c2, oc.qn, oc.oQ)
    for i, q in enumerate(Q):
        for k in range(0, oc.n, 61):
            a1, b1, a2, b2 = (int(v[i, k]) for v in (c1[0], c1[1], c2[0], c2[1]))
            assert int(out[0, i, k]) == a1 * a2 % q
            assert int(out[1, i, k]) == (a1 * b2 + a2 * b1) % q
            assert int(out[2, i, k]) == b1 * b2 % q


def test_delta_m_is_ceil_q_m_over_t(oc):
    rng = np.random.default_rng(6)
    Q = [int(x) for x in oc.mod[oc.oQ:oc.oQ + oc.qn]]
    Qp = 1
    for q in Q:
        Qp *= q
    m = rng.integers(0, oc.plain, size=oc.n, dtype=np.uint64)
    m[:3] = [0, 1, oc.plain - 1]
    out = oc.delta_m(m)
    t = oc.plain
    for k in range(0, oc.n, 17):
        want = -(-Qp * int(m[k]) // t)  # ceil
        assert [int(v) for v in out[:, k]] == [want % q for q in Q]
