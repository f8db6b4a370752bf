"""The exact pins of the oracle's estimate stages reject plausible mistakes
(CPU only).  Each case compiles the oracle sources with one planted mistake
(-DOR_MUTATION=n, see oracle/ctx_internal.h) and runs the same check that the
real build passes in test_oracle_hps.py / test_oracle_ks.py; the check must
fail on the planted build.

  1  scale-and-round without the rounding half (nu starts at 0, not 0.5)
  2  Q_j -> T and PQ -> T with the float estimate rounded, not truncated
  3  K > 1 modDown with an alpha * P correction the reference does not make
  4  scale-and-round with nu rounded instead of truncated
  5  modulus switching with the centred residue of the dropped limb (rounding
     instead of the reference's floor)
"""
import os
import subprocess

import pytest

import oracle_lib as orc
import params as Pm
import test_oracle_hps as th
import test_oracle_ks as tk
import test_oracle_levels as tl

SRC = ["nt.c", "ntt.c", "ctx.c", "hps.c", "ks.c", "xsalsa20.c", "bfv.c"]
ODIR = os.path.join(orc.ROOT, "oracle")


@pytest.fixture(scope="module")
def variants(tmp_path_factory):
    d = tmp_path_factory.mktemp("oracle_mut")
    procs = {}
    for n in (1, 2, 3, 4, 5):
        out = str(d / f"liboracle_mut{n}.so")
        cmd = ["gcc", "-O2", "-std=gnu11", "-fPIC", "-ffp-contract=off", f"-DOR_MUTATION={n}", "-shared",
               "-o", out] + [os.path.join(ODIR, f) for f in SRC] + ["-lm"]
        procs[n] = (subprocess.Popen(cmd), out)
    libs = {}
    for n, (p, out) in procs.items():
        assert p.wait() == 0, f"planted build {n} failed"
        libs[n] = orc.load_variant(out)
    return libs


def ctx_k1(lib=None):
    q, p, t = orc.chain_60bit(4, 1, 3)
    return orc.Ctx(12, q, p, t, 1, 65537, _lib=lib)


def ctx_ref14(lib=None):
    d = Pm.REF14
    return orc.Ctx(12, d["q"], d["p"], d["t"], d["gamma"], d["plain"], _lib=lib)


def test_real_build_passes_every_check():
    th.check_scale_round_exact_small(ctx_k1())
    for oc in (ctx_k1(), ctx_ref14()):
        tk.check_modup_to_t(oc, "qj", 1)
        tk.check_modup_to_t(oc, "pq", 2)
        tk.check_moddown(oc)
    tl.check_mod_switch_floor(tl.ctx_k1_n4096(), 3, 0)


@pytest.mark.parametrize("n", [1, 4])
def test_scale_round_pin_rejects_planted_mistake(variants, n):
    with pytest.raises(AssertionError):
        th.check_scale_round_exact_small(ctx_k1(variants[n]))


@pytest.mark.parametrize("which", ["qj", "pq"])
@pytest.mark.parametrize("mk", [ctx_k1, ctx_ref14])
def test_modup_pin_rejects_rounded_estimate(variants, which, mk):
    with pytest.raises(AssertionError):
        tk.check_modup_to_t(mk(variants[2]), which, 1 if which == "qj" else 2)


def test_moddown_pin_rejects_alpha_correction(variants):
    with pytest.raises(AssertionError):
        tk.check_moddown(ctx_ref14(variants[3]))


def test_mod_switch_pin_rejects_rounding_switch(variants):
    with pytest.raises(AssertionError):
        tl.check_mod_switch_floor(tl.ctx_k1_n4096(variants[5]), 3, 0)
